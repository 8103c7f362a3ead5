// capi.cpp — the C ABI of include/mps_sph.h on top of the gfx950 kernels.
//
// One context = one HIP stream + all device buffers of S >= 1 independent scenes stored
// back to back (one scene unless the batch entry points define several). The step runs the
// phases of System::OnUpdate (MPSSystem.cpp:191-291) in order, each phase as one launch
// over every scene. Solver loops are driven
// from the host in chunks: a chunk of iterations is enqueued back to back, the kernels
// keep the loop state on the device (SolverDev) and exit at once after convergence, and
// the host reads the 8-byte loop header once per chunk. The chunk length follows the
// previous step's iteration count, so a converging solve usually costs one host wait.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mps_sph.h"
#include "device_common.hpp"
#include "kernels.hpp"

namespace mps {
namespace launch {
thread_local Timing* timing = nullptr;
}
}  // namespace mps

namespace {

using mps::HashView;
using mps::NbrCounters;
using mps::SolverDev;
using mps::StepParams;
using mps::dev::HashGeom;

struct HipError {
    hipError_t e;
    const char* what;
};
#define HIPCHK(x)                                         \
    do {                                                  \
        hipError_t e_ = (x);                              \
        if (e_ != hipSuccess) throw HipError{e_, #x};     \
    } while (0)

struct ApiError {
    int code;
    std::string msg;
};

// Set while mps_step_noalloc runs on this thread: any buffer that would have to grow throws
// instead of allocating.
thread_local bool g_no_alloc = false;
[[noreturn]] void refuse_alloc() {
    throw ApiError{MPS_ERR_STATE,
                   "mps_step_noalloc: the context's buffers are smaller than this step needs (its neighbour lists grew, "
                   "or no mps_step has sized it yet); the state is the one before the step, and one mps_step continues "
                   "from it"};
}

template <class T>
struct DBuf {
    T* p = nullptr;
    size_t cap = 0;
    void reserve(size_t n) {
        if (n <= cap) return;
        if (g_no_alloc) refuse_alloc();
        if (p) HIPCHK(hipFree(p));
        p = nullptr;
        cap = 0;
        if (n) HIPCHK(hipMalloc(reinterpret_cast<void**>(&p), n * sizeof(T)));
        cap = n;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

struct HashSet {
    DBuf<uint32_t> key, id, key2, id2, start, end, temp;
    DBuf<float4> spos;
    uint32_t n_valid = 0;  // length of the sorted key array the cell table was built from
    bool cells_init = false;
    void release() {
        key.release(); id.release(); key2.release(); id2.release(); start.release(); end.release(); temp.release();
        spos.release();
    }
};

constexpr size_t al256(size_t b) { return (b + 255) & ~size_t(255); }

// ---- the allocation-free step's scratch
// mps_step_noalloc keeps the step's reduction scratch in this array instead of the context's
// buffers: per-workgroup FP64 partial sums, per-workgroup neighbour-count records, the
// neighbour counters, the three solver records of every scene and the running-scene counters.
// Sized for max(fluid slots, wall particles) <= kNoAllocSlots and S <= kNoAllocScenes (1.5 MiB);
// each GPU the library runs on holds its own copy until the process ends. Every call on a GPU
// uses that GPU's copy: calls that overlap on one GPU race.
constexpr size_t kNoAllocSlots = size_t(1) << 23;
constexpr uint32_t kNoAllocScenes = 64;
constexpr size_t kNoAllocBlocks = kNoAllocSlots / mps::dev::kBlock;
constexpr size_t kScrPartials = 3 * kNoAllocBlocks + 8;  // doubles (ensure_partials)
constexpr size_t kScrNstat = 2 * kNoAllocBlocks + 2;     // uint2 (neighbors)
constexpr size_t kScrOffNstat = al256(kScrPartials * sizeof(double));
constexpr size_t kScrOffCtr = kScrOffNstat + al256(kScrNstat * sizeof(uint2));
constexpr size_t kScrOffSolver = kScrOffCtr + al256(sizeof(NbrCounters));
constexpr size_t kScrOffActive = kScrOffSolver + al256(3 * size_t(kNoAllocScenes) * sizeof(SolverDev));
constexpr size_t kScrBytes = kScrOffActive + al256(4 * sizeof(uint32_t));
static_assert(kScrBytes <= (size_t(3) << 19), "scratch stays modest");

static __device__ __align__(256) unsigned char scratch[kScrBytes];

// The address of `scratch` on `device` (which must be current), looked up once per device.
// Host code uses only this address, never the array's name, which on the host is a stand-in.
unsigned char* scratch_address(int device) {
    static std::mutex mu;
    static std::vector<unsigned char*> addr;
    std::lock_guard<std::mutex> lock(mu);
    if (device < 0) throw ApiError{MPS_ERR_INVALID, "bad device"};
    if (addr.size() <= size_t(device)) addr.resize(size_t(device) + 1, nullptr);
    if (!addr[device]) {
        void* ptr = nullptr;
        HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) throw ApiError{MPS_ERR_HIP, "no device address for the no-alloc scratch"};
        addr[device] = static_cast<unsigned char*>(ptr);
    }
    return addr[device];
}

int bits_for(size_t count) {  // number of bits needed for values < count
    int b = 0;
    while ((size_t(1) << b) < count) b++;
    return std::max(b, 1);
}

}  // namespace

struct mps_ctx {
    mps_config cfg{};
    mps_physics phys{};
    mps_sph_material mat{};
    double mV = 0.0, mMass = 0.0;
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;

    // N = fluid slots (scenes back to back, each starting on a kBlock boundary), Nref = fluid
    // particles (reference ids 0..Nref), B = boundary particles. With one scene N == Nref.
    uint32_t N = 0, Nref = 0, B = 0, frame = 0;
    // largest fluid slots, wall particles and scenes a completed step has sized the buffers for
    uint32_t sized_N = 0, sized_B = 0, sized_S = 0;
    uint32_t zsorted_N = 0, zsorted_B = 0;  // the same for the Z-sort's temporaries

    // ------------------------------------------------------------------ scenes
    struct SceneHost {
        mps_physics phys{};
        mps_sph_material mat{};
        double V = 0.0, mass = 0.0;
        uint32_t nf = 0, nb = 0;  // fluid / boundary particles
        uint32_t begin = 0;       // first fluid slot
        uint32_t ref0 = 0;        // first reference id
        uint32_t b0 = 0;          // first boundary particle
    };
    struct StagedScene {
        SceneHost h;
        std::vector<float> pos, vel, rad, mass, bpos, brad, bmass;
    };
    std::vector<SceneHost> scenes;      // batch layout; empty = one scene made of everything uploaded
    std::vector<StagedScene> staged;    // scenes added since the last commit
    std::vector<mps_step_stats> sstats; // per-scene statistics of the last step
    uint32_t S = 1;
    DBuf<StepParams> d_params;
    DBuf<uint32_t> blk_scene, bscene, d_delta, mkey;
    bool params_dirty = true;

    // ------------------------------------------------------------------ adjoint state
    // One recorded step: what its reverse needs (positions, frozen lists, walls, the velocity
    // after gravity and after every DF/CD force and the CG solve, per-scene loop counts).
    struct TapeStep {
        uint32_t Kff = 0, Kfb = 0;
        DBuf<float4> pos, bpos;
        DBuf<float> bvol;
        DBuf<uint32_t> src, nff, cff, nfb, cfb, perm_end;
        std::vector<DBuf<float4>> snaps;
        uint32_t nsnap = 0, df_last = 0, cg_out = 0;
        std::vector<uint32_t> ldf, lcd;
        void release() {
            pos.release(); bpos.release(); bvol.release(); src.release(); nff.release(); cff.release();
            nfb.release(); cfb.release(); perm_end.release();
            for (auto& b : snaps) b.release();
        }
    };
    // Everything a step reads that earlier steps wrote.
    struct Checkpoint {
        DBuf<float4> pos, vel, bpos, bvel, bcolor;
        DBuf<float> bvol;
        DBuf<uint32_t> perm, bface;
        DBuf<float2> bbcc;
        uint32_t frame = 0;
        void release() {
            pos.release(); vel.release(); bpos.release(); bvel.release(); bcolor.release(); bvol.release();
            perm.release(); bface.release(); bbcc.release();
        }
    };
    std::vector<TapeStep> tape;
    TapeStep* rec = nullptr;
    std::vector<Checkpoint> ckpts;
    Checkpoint ck_end;
    uint32_t adj_n = 0, adj_every = 0;
    // layout the rollout was recorded on: a backward on any other layout is refused
    uint32_t adj_N = 0, adj_Nref = 0, adj_B = 0, adj_S = 0;
    bool adj_eq = true;  // every scene's fluid radii equal (the reverse kernels' fast path)
    // Forget the recorded rollout: called by everything that replaces particles, walls, scenes
    // or the parameters the recomputation would read.
    void drop_rollout() {
        adj_n = 0;
        lam_valid = false;
    }
    DBuf<uint32_t> perm0, d_L, rev;
    DBuf<float> gobs;  // one observation's cotangents (Nref x 3 position, then velocity)
    DBuf<float4> adx, adv, adx2, adv2, aux4, gvec, avin, lam, lam2;
    bool lam_valid = false;  // lam holds the later step's transposed solution (warm start)
    DBuf<float> abar, rbar, dbar, fbar, fstr;
    DBuf<float2> cabT;
    DBuf<double> acc_mu, acc_sig;
    // fluid: pos = (x, y, z, r), vel = (vx, vy, vz, m)
    DBuf<float4> pos, vel, color, cgb, cgx, cgr, cgAp, jprev, jt[3], jcand, tmp4;
    // the forward CG's vectors as packed triples: x, r, A p and two p buffers
    DBuf<mps::dev::f3> v3x, v3r, v3Ap, v3p, v3q;
    DBuf<float> rho, alpha, kappa, delta, fst, tmp1;
    // second buffer of each array the hash reorders: the two alternate, so the arrays a solver
    // chunk reads take one of two addresses (cached solver graphs stay valid)
    DBuf<float4> pos_alt, vel_alt, color_alt, cgb_alt;
    DBuf<float> rho_alt, alpha_alt, kappa_alt, delta_alt, fst_alt;
    // boundary: bpos = (x, y, z, r), bvel = (vx, vy, vz, m)
    DBuf<float4> bpos, bvel, bcolor;
    DBuf<float> bvol, bS;
    DBuf<uint32_t> bface;
    DBuf<float2> bbcc;
    // The mesh of the last mps_sample_boundary_mesh (vertices after fitting, faces, ownership
    // bits, vertex masses) and each face's sample count and shortest edge, as MeshObject keeps
    // them between ParticleSampling calls (m_samplingParticleSize, m_shortestEdgeID).
    struct MeshKeep {
        std::vector<double> V, vmass;
        std::vector<uint32_t> F, R, count, shorts;
        bool ok = false;
    } mesh;
    bool bnd_dirty = true;
    HashSet fh, bh;
    DBuf<uint32_t> zkey, zid, zkey2, zid2, ztemp, tmpu;
    DBuf<float2> tmp2;
    // The fluid's working arrays are kept in slot order (sorted by cell key, ties by
    // reference index). perm[s] = reference index of slot s, inv = its inverse.
    DBuf<uint32_t> perm, inv, keyref, vals, src, lkey, hash_perm;
    // neighbour lists (column-major ELL)
    DBuf<uint32_t> nff, cff, nfb, cfb, nbb, cbb;
    uint32_t Kff = 0, Kfb = 0, Kbb = 0;
    // coupled fluid objects: ref lists into the other scenes (column-major, Kfr slots)
    DBuf<uint32_t> nfr, cfr;
    uint32_t Kfr = 0;
    // wall objects (mps_set_boundary_objects): object id per wall particle, per-object sums
    uint32_t nbobj = 1;
    DBuf<uint32_t> bobj;
    DBuf<float4> bS4;
    bool coupled() const { return cfg.couple_scenes != 0 && S > 1; }
    // per neighbour slot: G_ij (ff, fb) for the pressure solves, a_ij / b_ij for the CG operator
    DBuf<float> pairG_ff, pairG_fb;
    DBuf<float2> pairAB;  // (a_ij, b_ij) of the implicit viscosity operator
    DBuf<float> pairA;    // a_ij alone, when A.p recomputes b_ij (cfg.cg_stream_coeffs == 0)
    DBuf<float> irho;     // 1/rho per slot for that recomputation
    // block halos of the fluid lists for A.p (halo.hip): sorted distinct neighbour slots per
    // workgroup, their count, and every list entry as a 16-bit position in its block's halo
    DBuf<uint32_t> halo_idx, halo_n;
    DBuf<uint16_t> nb16;
    DBuf<uint2> nstat;  // per-workgroup neighbour-count records of the searches
    DBuf<float4> cgrecA, cgrecB;  // 32-byte (position, vector) records for the CG gathers
    bool pairs_valid = false;
    bool halo_valid = false;  // block halos match the fluid lists
    uint32_t nei_regrow = 0;
    // small device records + pinned host mirrors
    NbrCounters* d_ctr = nullptr;
    NbrCounters* h_ctr = nullptr;
    SolverDev* d_solver = nullptr;  // [3][S]: DF, CD, CG records of every scene
    SolverDev* h_solver = nullptr;
    uint32_t solver_cap = 0;        // scenes the two arrays above hold
    uint32_t* d_active = nullptr;   // [3]: scenes whose DF / CD / CG loop is still running
    uint32_t* h_active = nullptr;
    mps::GridSync* d_grid = nullptr;  // barrier words of the persistent solver (allocated at creation)
    uint32_t persist_cap[6] = {0, 0, 0, 0, 0, 0};  // co-resident workgroups per kernel variant (0 = not asked yet)
    uint32_t persist_fallbacks = 0;       // persistent solves that gave up and reran chunked
    uint32_t deep_halos = 0;              // workgroups of the last list build with halos over kHaloShallow
    uint32_t* d_flags = nullptr;  // [0] nonfinite, [1] sampling mismatch, [4..6] halo statistics of the step
    uint32_t* h_flags = nullptr;
    double* d_err3 = nullptr;
    double* h_err3 = nullptr;
    DBuf<double> partials;
    DBuf<uint32_t> arrive;  // per-workgroup arrival counters of the CG launches that fold in their last workgroup
    // ---- render feed: the state is packed on the compute stream into one of two staging slots
    // and copied to the host on a second stream behind an event, so the next step's kernels run
    // while the copy drains; a slot is packed again only after its previous copy finished.
    struct ReadbackSlot {
        DBuf<unsigned char> buf;
        hipEvent_t packed = nullptr, copied = nullptr;
        bool used = false;
    } rb[2];
    hipStream_t copy_stream = nullptr;
    uint64_t rb_seq = 0;
    void readback(void* host_pos, void* host_radius, float* host_rgba, bool f64, mps_fence* fence) {
        if (!copy_stream) HIPCHK(hipStreamCreateWithFlags(&copy_stream, hipStreamNonBlocking));
        ReadbackSlot& sl = rb[rb_seq & 1u];
        if (!sl.packed) HIPCHK(hipEventCreateWithFlags(&sl.packed, hipEventDisableTiming));
        if (!sl.copied) HIPCHK(hipEventCreateWithFlags(&sl.copied, hipEventDisableTiming));
        const size_t n = Nref, m = std::max<size_t>(n, 1);
        const size_t real = f64 ? sizeof(double) : sizeof(float);
        const size_t o_rad = al256(3 * m * real), o_rgba = o_rad + al256(m * real);
        const size_t bytes = o_rgba + m * sizeof(float4);
        if (sl.used && bytes > sl.buf.cap) HIPCHK(hipEventSynchronize(sl.copied));  // no free under a copy
        sl.buf.reserve(bytes);
        if (sl.used) HIPCHK(hipStreamWaitEvent(stream, sl.copied, 0));  // the slot's last copy has drained
        unsigned char* b = sl.buf.p;
        if (f64)
            mps::launch::pack_readback_f64(N, pos.p, color.p, perm.p, host_pos ? reinterpret_cast<double*>(b) : nullptr,
                                           host_radius ? reinterpret_cast<double*>(b + o_rad) : nullptr,
                                           host_rgba ? reinterpret_cast<float*>(b + o_rgba) : nullptr, stream);
        else
            mps::launch::pack_readback(N, pos.p, color.p, perm.p, host_pos ? reinterpret_cast<float*>(b) : nullptr,
                                       host_radius ? reinterpret_cast<float*>(b + o_rad) : nullptr,
                                       host_rgba ? reinterpret_cast<float*>(b + o_rgba) : nullptr, stream);
        HIPCHK(hipEventRecord(sl.packed, stream));
        HIPCHK(hipStreamWaitEvent(copy_stream, sl.packed, 0));
        if (n) {
            if (host_pos) HIPCHK(hipMemcpyAsync(host_pos, b, 3 * n * real, hipMemcpyDeviceToHost, copy_stream));
            if (host_radius) HIPCHK(hipMemcpyAsync(host_radius, b + o_rad, n * real, hipMemcpyDeviceToHost, copy_stream));
            if (host_rgba)
                HIPCHK(hipMemcpyAsync(host_rgba, b + o_rgba, n * sizeof(float4), hipMemcpyDeviceToHost, copy_stream));
        }
        HIPCHK(hipEventRecord(sl.copied, copy_stream));
        sl.used = true;
        fence->event = sl.copied;
        fence->frame = frame;
        rb_seq++;
    }
    uint32_t pred[3] = {2, 3, 8};
    uint32_t pred_adj = 8;  // iterations of the last transposed solve of the adjoint
    // Iterations enqueued before the next host poll: the predicted count first (the previous
    // solve's), then small chunks, so a solve that runs a little longer than predicted does not
    // enqueue a whole second prediction of launches that would exit at once.
    static constexpr uint32_t kFollowChunk = 2;
    uint32_t chunk_len(uint32_t it, uint32_t predicted, uint32_t left) const {
        if (cfg.host_sync_every) return std::min(cfg.host_sync_every, left);
        return std::min(it == 1 ? std::max<uint32_t>(predicted, 1) : kFollowChunk, left);
    }

    // ------------------------------------------------------------------ kernel timing
    enum KId {
        K_KEYS, K_SORT, K_CELLS, K_NEI_FLUID, K_NEI_BND, K_BSUM, K_BVOL, K_DENSITY, K_EVAL_DF, K_EVAL_CD,
        K_APPLY, K_ST_FACTOR, K_CG_SETUP, K_CG_INIT, K_CG_APPLY, K_CG_UPDATE, K_CG_DIR, K_COPY_OUT, K_ADVECT,
        K_ZSORT, K_JACOBI, K_FOLD, K_REORDER, K_HALO, K_CALIB, K_STREAM, K_COUNT
    };
    static const char* kname(int k) {
        static const char* names[K_COUNT] = {
            "k_cell_keys", "radix_sort", "k_cell_table", "k_neighbors_fluid", "k_neighbors_boundary",
            "k_boundary_sum", "k_boundary_volume", "k_density_alpha", "k_pressure_eval<DF>", "k_pressure_eval<CD>",
            "k_pressure_apply", "k_st_factor", "k_cg_setup", "k_cg_apply<0>", "k_cg_apply<1>", "k_cg_update",
            "k_cg_direction", "k_copy_vel_out", "k_advect", "zsort", "jacobi", "k_fold", "reorder", "k_halo_build",
            "k_calib_copy", "k_stream_copy"};
        return names[k];
    }
    struct Rec {
        int k;
        hipEvent_t a, b;
    };
    bool prof_on = false;
    std::vector<Rec> recs;
    std::vector<hipEvent_t> evpool;
    double kms[K_COUNT] = {};
    uint64_t kn[K_COUNT] = {}, kuse[K_COUNT] = {};
    hipEvent_t ev_get() {
        if (!evpool.empty()) {
            hipEvent_t e = evpool.back();
            evpool.pop_back();
            return e;
        }
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        return e;
    }
    // Hands every kernel launched inside T_ its own event pair (stamped by the dispatch itself).
    struct KTimer final : mps::launch::Timing {
        mps_ctx* c = nullptr;
        int k = 0;
        uint32_t n = 0;
        void events(hipEvent_t* a, hipEvent_t* b) override {
            *a = c->ev_get();
            *b = c->ev_get();
            c->recs.push_back(Rec{k, *a, *b});
            n++;
        }
    } ktimer;
    // Run launch(s) f; when profiling, time each of their kernels (durations add up per id k).
    template <class F>
    void T_(int k, F&& f, bool always_useful = true) {
        if (!prof_on) {
            f();
            return;
        }
        ktimer.c = this;
        ktimer.k = k;
        ktimer.n = 0;
        struct Scope {
            explicit Scope(mps::launch::Timing* t) { mps::launch::timing = t; }
            ~Scope() { mps::launch::timing = nullptr; }
        } scope(&ktimer);
        f();
        if (ktimer.n) {
            kn[k]++;
            if (always_useful) kuse[k]++;
        }
    }
    void prof_collect() {
        for (const Rec& r : recs) {
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, r.a, r.b));
            kms[r.k] += ms;
            evpool.push_back(r.a);
            evpool.push_back(r.b);
        }
        recs.clear();
    }
    void prof_reset() {
        prof_collect();
        for (int k = 0; k < K_COUNT; k++) { kms[k] = 0.0; kn[k] = 0; kuse[k] = 0; }
    }

    // ------------------------------------------------------------------ graph replay
    // A solver chunk is a fixed sequence of launches whose arguments do not depend on the
    // iteration, so it is captured once per argument set and replayed.
    struct GraphEntry {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
    };
    std::map<std::vector<uint64_t>, GraphEntry> graphs;
    bool graphs_on() const { return cfg.use_graphs && !prof_on && !rec && !g_no_alloc; }
    void clear_graphs() {
        for (auto& kv : graphs) {
            if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
            if (kv.second.graph) (void)hipGraphDestroy(kv.second.graph);
        }
        graphs.clear();
    }
    template <class F>
    void launch_chunk(std::vector<uint64_t> key, F&& f) {
        if (!graphs_on()) {
            f();
            return;
        }
        auto it = graphs.find(key);
        if (it == graphs.end()) {
            if (graphs.size() >= 64) clear_graphs();
            GraphEntry e;
            HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
            try {
                f();
            } catch (...) {  // leave the stream out of capture mode before reporting
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(stream, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            HIPCHK(hipStreamEndCapture(stream, &e.graph));
            HIPCHK(hipGraphInstantiate(&e.exec, e.graph, nullptr, nullptr, 0));
            it = graphs.emplace(std::move(key), e).first;
        }
        HIPCHK(hipGraphLaunch(it->second.exec, stream));
    }
    static uint64_t u(const void* p) { return uint64_t(reinterpret_cast<uintptr_t>(p)); }
    static uint64_t u(double d) {
        uint64_t b;
        std::memcpy(&b, &d, sizeof(b));
        return b;
    }
    std::vector<uint64_t> key_base(uint64_t kind, uint32_t chunk, const mps::Scenes& P, const mps::Nbr& nb) const {
        return {kind, chunk, u(P.P), u(P.blk_scene), P.N, P.S, u(nb.ff), u(nb.cff), u(nb.fb), u(nb.cfb), u(nb.fr), u(nb.cfr), u(nb.fbo), nb.nbo};
    }

    // ------------------------------------------------------------------ helpers
    HashGeom geom() const {
        const double cell = cfg.cell_size > 0.0 ? double(float(cfg.cell_size)) : double(float(mat.radius));
        HashGeom g;
        g.hx = cfg.hash_size[0];
        g.hy = cfg.hash_size[1];
        g.hz = cfg.hash_size[2];
        g.inv_cell = 1.0 / cell;
        return g;
    }
    size_t cells() const { return size_t(cfg.hash_size[0]) * cfg.hash_size[1] * cfg.hash_size[2]; }
    // The scene table: the batch layout, or one scene covering everything uploaded.
    std::vector<SceneHost> scene_list() const {
        if (!scenes.empty()) return scenes;
        SceneHost h;
        h.phys = phys;
        h.mat = mat;
        h.V = mV;
        h.mass = mMass;
        h.nf = N;
        h.nb = B;
        return {h};
    }
    StepParams params_of(const SceneHost& h) const {
        StepParams P{};
        P.N = N;
        P.begin = h.begin;
        P.end = h.begin + h.nf;
        P.count = h.nf;
        P.blk0 = h.begin / mps::dev::kBlock;
        P.nblk = (h.nf + mps::dev::kBlock - 1) / mps::dev::kBlock;
        P.V = float(h.V);
        P.rho0 = float(h.mat.density);
        P.mass = float(h.mass);
        P.mu = float(h.mat.viscosity);
        P.xsph = float(h.mat.xsph);
        P.sigma = float(h.mat.surface_tension);
        P.dt = float(h.phys.dt);
        P.gx = float(h.phys.gravity[0]);
        P.gy = float(h.phys.gravity[1]);
        P.gz = float(h.phys.gravity[2]);
        for (int k = 0; k < 3; k++) {
            P.bmin[k] = float(h.phys.box_min[k]);
            P.bmax[k] = float(h.phys.box_max[k]);
        }
        P.cg_factor = float(cfg.cg_factor);
        P.faithful = cfg.faithful_quirks;
        return P;
    }
    // Upload the per-scene parameters (and the workgroup -> scene map for one scene) when
    // something they depend on changed.
    void sync_params() {
        if (!params_dirty) return;
        const std::vector<SceneHost> sl = scene_list();
        S = uint32_t(sl.size());
        std::vector<StepParams> hp;
        for (const SceneHost& h : sl) hp.push_back(params_of(h));
        d_params.reserve(hp.size());
        HIPCHK(hipMemcpyAsync(d_params.p, hp.data(), hp.size() * sizeof(StepParams), hipMemcpyHostToDevice, stream));
        if (scenes.empty()) {
            const size_t nblk = std::max<size_t>((size_t(N) + mps::dev::kBlock - 1) / mps::dev::kBlock, 1);
            blk_scene.reserve(nblk);
            mps::launch::zero_words(blk_scene.p, nblk, stream);
            d_delta.reserve(1);
            mps::launch::zero_words(d_delta.p, 1, stream);
        }
        ensure_solvers();
        sync();
        params_dirty = false;
    }
    void ensure_solvers() {
        if (solver_cap >= S) return;
        if (g_no_alloc) refuse_alloc();
        if (d_solver) HIPCHK(hipFree(d_solver));
        if (h_solver) HIPCHK(hipHostFree(h_solver));
        d_solver = nullptr;
        h_solver = nullptr;
        solver_cap = 0;
        HIPCHK(hipMalloc(reinterpret_cast<void**>(&d_solver), 3 * size_t(S) * sizeof(SolverDev)));
        HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&h_solver), 3 * size_t(S) * sizeof(SolverDev)));
        solver_cap = S;
        for (int k = 0; k < 3; k++) mps::launch::solver_reset(d_solver + size_t(k) * S, S, d_active + k, stream);
    }
    mps::Scenes scn() const { return mps::Scenes{d_params.p, blk_scene.p, N, S}; }
    // bits of the scene index placed above the cell bits of a sort key
    int scene_bits() const { return S > 1 ? bits_for(S) : 0; }
    int morton_bits() const {
        const uint32_t hmax = std::max(cfg.hash_size[0], std::max(cfg.hash_size[1], cfg.hash_size[2]));
        return 3 * std::min(bits_for(hmax), 10);
    }
    int cell_bits() const { return bits_for(cells()); }
    const uint32_t* bnd_scene() const { return S > 1 ? bscene.p : nullptr; }
    mps::Nbr nbr() const {
        mps::Nbr nb{nff.p, cff.p, nfb.p, cfb.p};
        if (coupled()) {
            nb.fr = nfr.p;
            nb.cfr = cfr.p;
        }
        if (nbobj > 1) {
            nb.fbo = bobj.p;
            nb.nbo = nbobj;
        }
        return nb;
    }
    int cubic() const { return cfg.kernel == MPS_KERNEL_CUBIC; }
    void sync() { HIPCHK(hipStreamSynchronize(stream)); }

    void ensure_partials() {
        const size_t nblk = (std::max<size_t>(std::max<size_t>(N, B), 1) + mps::dev::kBlock - 1) / mps::dev::kBlock;
        partials.reserve(3 * nblk + 8);
    }
    // The arrival counters of the fold-in-kernel CG launches: zero when (re)allocated; every
    // launch leaves them at zero.
    uint32_t* arrive_counters() {
        const size_t nblk = (std::max<size_t>(N, 1) + mps::dev::kBlock - 1) / mps::dev::kBlock;
        if (arrive.cap < nblk) {
            arrive.reserve(nblk);
            mps::launch::zero_words(arrive.p, arrive.cap, stream);
        }
        return arrive.p;
    }
    size_t table_cells() const { return size_t(S) * cells(); }

    // ------------------------------------------------------------------ phases
    void zsort_set(uint32_t n, DBuf<float4>& p4, bool boundary) {
        if (n == 0) return;
        const HashGeom g = geom();
        zkey.reserve(n); zid.reserve(n); zkey2.reserve(n); zid2.reserve(n);
        ztemp.reserve(mps::launch::radix_temp_words(n));
        const int zb = morton_bits();
        // several wall objects: each is Z-sorted within its own range (the object id above the
        // Morton key), as each reference object Z-sorts with its own hash
        const bool objs = boundary && nbobj > 1;
        const uint32_t* zscene = objs ? bobj.p : bnd_scene();
        // every temporary first, so nothing is permuted unless the whole sort can run
        tmp4.reserve(n);
        if (boundary) {
            tmpu.reserve(n);
            tmp2.reserve(n);
            if (!cfg.faithful_quirks) tmp1.reserve(n);
        }
        T_(K_ZSORT, [&] { mps::launch::cell_keys(p4.p, n, g, 1, zscene, zb, zkey.p, zid.p, stream); });
        const int bits = zb + (objs ? bits_for(nbobj) : scene_bits());
        uint32_t *k = zkey.p, *v = zid.p, *k2 = zkey2.p, *v2 = zid2.p;
        T_(K_ZSORT, [&] { mps::launch::radix_sort(k, v, k2, v2, ztemp.p, n, bits, stream); });
        const uint32_t* perm = v;
        auto perm4 = [&](DBuf<float4>& a) {
            T_(K_ZSORT, [&] { mps::launch::gather_f4(perm, n, a.p, tmp4.p, stream); });
            std::swap(a.p, tmp4.p);
            std::swap(a.cap, tmp4.cap);
        };
        perm4(p4);
        if (boundary) {
            perm4(bvel);
            perm4(bcolor);
            mps::launch::gather_u32(perm, n, bface.p, tmpu.p, stream);
            std::swap(bface.p, tmpu.p); std::swap(bface.cap, tmpu.cap);
            mps::launch::gather_f2(perm, n, bbcc.p, tmp2.p, stream);
            std::swap(bbcc.p, tmp2.p); std::swap(bbcc.cap, tmp2.cap);
            if (!cfg.faithful_quirks) {
                mps::launch::gather_f1(perm, n, bvol.p, tmp1.p, stream);
                std::swap(bvol.p, tmp1.p); std::swap(bvol.cap, tmp1.cap);
            }
            bnd_dirty = true;
        } else {
            perm4(vel);
            perm4(color);
        }
    }
    // Fluid Z-sort: the reference permutes its arrays by Morton key; here the data stays in
    // slot order and only the reference labels change (perm[s] <- new index of perm[s]).
    // Keys sort scene by scene, so scene s keeps the reference ids [ref0, ref0 + nf).
    void zsort_fluid() {
        if (N == 0) return;
        const HashGeom g = geom();
        keyref.reserve(N); vals.reserve(N); zkey2.reserve(N); zid2.reserve(N); zid.reserve(N);
        ztemp.reserve(mps::launch::radix_temp_words(N));
        const int zb = morton_bits();
        const int bits = zb + scene_bits();
        T_(K_ZSORT, [&] {
            mps::launch::cell_keys_slots(pos.p, perm.p, N, g, 1, blk_scene.p, zb, keyref.p, vals.p, stream);
            uint32_t *k = keyref.p, *v = vals.p, *k2 = zkey2.p, *v2 = zid2.p;
            mps::launch::radix_sort(k, v, k2, v2, ztemp.p, Nref, bits, stream);
            mps::launch::invert(v, Nref, zid.p, stream);    // zid[old ref] = new ref
            mps::launch::relabel(perm.p, zid.p, N, stream);
            mps::launch::invert(perm.p, N, inv.p, stream);
        });
    }
    void zsort() {
        zsort_set(B, bpos, true);
        zsort_fluid();
    }
    void hash_set(HashSet& h, uint32_t n, const float4* p4) {
        const size_t C = table_cells();
        h.start.reserve(C);
        h.end.reserve(C);
        if (!h.cells_init) {
            T_(K_CELLS, [&] {
                mps::launch::fill_u32(h.start.p, C, mps::dev::kEmpty, stream);
                mps::launch::fill_u32(h.end.p, C, mps::dev::kEmpty, stream);
            });
            h.cells_init = true;
        } else {
            T_(K_CELLS, [&] { mps::launch::clear_cells(h.key.p, h.n_valid, h.start.p, h.end.p, stream); });
        }
        h.n_valid = 0;
        if (n == 0) return;
        h.key.reserve(n); h.id.reserve(n); h.key2.reserve(n); h.id2.reserve(n); h.spos.reserve(n);
        h.temp.reserve(mps::launch::radix_temp_words(n));
        const int cb = cell_bits();
        T_(K_KEYS, [&] { mps::launch::cell_keys(p4, n, geom(), 0, bnd_scene(), cb, h.key.p, h.id.p, stream); });
        T_(K_SORT, [&] { mps::launch::radix_sort(h.key.p, h.id.p, h.key2.p, h.id2.p, h.temp.p, n, cb + scene_bits(), stream); });
        // key/key2 (id/id2) are always reserved together, so their capacities are equal and
        // the pointer swap inside radix_sort keeps each DBuf consistent.
        T_(K_CELLS, [&] { mps::launch::cell_table(h.key.p, h.id.p, n, h.start.p, h.end.p, p4, h.spos.p, stream); });
        h.n_valid = n;
    }
    // Fluid UpdateHash with slot-order storage. Slots are sorted by the Morton code of their
    // cell (values = reference ids, stable => ties by reference id), so every cell is a
    // contiguous slot range in ascending reference order — the candidate order of the
    // reference's sorted (key, id) table — while blocks of slots are compact in space. The
    // cell start/end table is indexed by the reference's linear key. The reference's
    // linear-order key/id arrays are rebuilt from lkey + perm only when exported.
    void hash_fluid() {
        HashSet& h = fh;
        const size_t C = table_cells();
        const HashGeom g = geom();
        h.start.reserve(C);
        h.end.reserve(C);
        if (!h.cells_init) {
            T_(K_CELLS, [&] {
                mps::launch::fill_u32(h.start.p, C, mps::dev::kEmpty, stream);
                mps::launch::fill_u32(h.end.p, C, mps::dev::kEmpty, stream);
            });
            h.cells_init = true;
        } else {
            T_(K_CELLS, [&] { mps::launch::clear_cells_m(mkey.p, lkey.p, h.n_valid, h.start.p, h.end.p, stream); });
        }
        h.n_valid = 0;
        if (N == 0) return;
        keyref.reserve(N); vals.reserve(N); src.reserve(N); lkey.reserve(N); hash_perm.reserve(N);
        h.key.reserve(N); h.id.reserve(N); h.key2.reserve(N); h.id2.reserve(N);
        h.temp.reserve(mps::launch::radix_temp_words(N));
        const int mb = morton_bits();
        T_(K_KEYS, [&] { mps::launch::cell_morton_slots(pos.p, perm.p, N, g, blk_scene.p, mb, keyref.p, vals.p, stream); });
        const int bits = mb + scene_bits();
        uint32_t *k = keyref.p, *v = vals.p, *k2 = h.key2.p, *v2 = h.id2.p;
        T_(K_SORT, [&] { mps::launch::radix_sort(k, v, k2, v2, h.temp.p, Nref, bits, stream); });
        if (k == keyref.p) {
            std::swap(h.key, keyref);
            std::swap(h.id, vals);
        } else {
            std::swap(h.key, h.key2);
            std::swap(h.id, h.id2);
        }
        T_(K_REORDER, [&] {
            // sorted position q -> slot; perm, src (old slot) and the key per slot
            mps::launch::place(h.key.p, h.id.p, Nref, mb, d_delta.p, inv.p, perm.p, src.p, mkey.p, stream);
            auto mv4 = [&](DBuf<float4>& a, DBuf<float4>& alt) {
                alt.reserve(N);
                mps::launch::gather_f4(src.p, N, a.p, alt.p, stream);
                std::swap(a, alt);
            };
            auto mv1 = [&](DBuf<float>& a, DBuf<float>& alt) {
                alt.reserve(N);
                mps::launch::gather_f1(src.p, N, a.p, alt.p, stream);
                std::swap(a, alt);
            };
            mv4(pos, pos_alt); mv4(vel, vel_alt); mv4(color, color_alt); mv4(cgb, cgb_alt);
            mv1(rho, rho_alt); mv1(alpha, alpha_alt); mv1(kappa, kappa_alt); mv1(delta, delta_alt); mv1(fst, fst_alt);
            HIPCHK(hipMemcpyAsync(hash_perm.p, h.id.p, sizeof(uint32_t) * Nref, hipMemcpyDeviceToDevice, stream));
            mps::launch::invert(perm.p, N, inv.p, stream);
        });
        if (rec) dcopy(rec->src, src.p, N);
        T_(K_CELLS, [&] {
            mps::launch::cell_table_m(mkey.p, pos.p, N, g, mb, uint32_t(cells()), h.start.p, h.end.p, lkey.p, stream);
        });
        h.n_valid = N;
        pairs_valid = false;
        halo_valid = false;
    }
    // The reference's fluid hash in its own layout (keys sorted by linear cell index, ties by
    // reference id; MPSSpatialHash.cu:53-77), rebuilt on the host from the last hash.
    void fluid_hash_export(std::vector<uint32_t>* keys, std::vector<uint32_t>* ids) {
        const uint32_t n = fh.n_valid;
        std::vector<uint32_t> lk = down(lkey.p, n), hp = down(hash_perm.p, n);
        std::vector<uint64_t> kv(n);
        for (uint32_t s = 0; s < n; s++) kv[s] = (uint64_t(lk[s]) << 32) | hp[s];
        std::sort(kv.begin(), kv.end());
        if (keys) { keys->resize(n); for (uint32_t s = 0; s < n; s++) (*keys)[s] = uint32_t(kv[s] >> 32); }
        if (ids) { ids->resize(n); for (uint32_t s = 0; s < n; s++) (*ids)[s] = uint32_t(kv[s]); }
    }
    std::vector<uint32_t> fluid_cells_export(bool want_start) {
        std::vector<uint32_t> keys;
        fluid_hash_export(&keys, nullptr);
        std::vector<uint32_t> out(fh.cells_init ? cells() : 0, mps::dev::kEmpty);
        const uint32_t n = uint32_t(keys.size());
        for (uint32_t s = 0; s < n; s++) {
            if (s == 0 || keys[s - 1] != keys[s]) {
                if (want_start) out[keys[s]] = s;
                else if (s > 0) out[keys[s - 1]] = s;
            }
            if (s == n - 1 && !want_start) out[keys[s]] = n;
        }
        return out;
    }
    void hash() {
        if (bnd_dirty || !bh.cells_init) hash_set(bh, B, bpos.p);
        hash_fluid();
    }
    HashView view(const HashSet& h, uint32_t n) const { return HashView{h.start.p, h.end.p, h.id.p, h.spos.p, n}; }
    HashView fluid_view() const { return HashView{fh.start.p, fh.end.p, nullptr, pos.p, N}; }

    void reset_counters() { mps::launch::zero_words(reinterpret_cast<uint32_t*>(d_ctr), sizeof(NbrCounters) / 4, stream); }

    void neighbors() {
        const HashGeom g = geom();
        // capacities rounded to a multiple of 4
        if (Kff == 0) Kff = (std::max<uint32_t>(cfg.nei_capacity, 8) + 3u) & ~3u;
        if (Kfb == 0) Kfb = (std::max<uint32_t>(cfg.nei_capacity / 2, 8) + 3u) & ~3u;
        if (Kbb == 0) Kbb = (std::max<uint32_t>(cfg.nei_capacity / 2, 8) + 3u) & ~3u;
        if (Kfr == 0) Kfr = (std::max<uint32_t>(cfg.nei_capacity / 2, 8) + 3u) & ~3u;
        for (int attempt = 0; attempt < 8; attempt++) {
            reset_counters();
            nstat.reserve(2 * ((size_t(std::max(N, B)) + mps::dev::kBlock - 1) / mps::dev::kBlock) + 2);
            const bool do_b = bnd_dirty;
            if (do_b && B) {
                nbb.reserve(size_t(Kbb) * B);
                cbb.reserve(B);
                T_(K_NEI_BND, [&] {
                    mps::launch::neighbors_boundary(bpos.p, B, bnd_scene(), uint32_t(cells()), g, view(bh, B), nbb.p, Kbb, cbb.p,
                                                    nstat.p, d_ctr, stream);
                });
            }
            nff.reserve(size_t(Kff) * std::max<uint32_t>(N, 1));
            cff.reserve(std::max<uint32_t>(N, 1));
            nfb.reserve(size_t(Kfb) * std::max<uint32_t>(N, 1));
            cfb.reserve(std::max<uint32_t>(N, 1));
            T_(K_NEI_FLUID, [&] {
                mps::launch::neighbors_fluid(pos.p, scn(), uint32_t(cells()), g, fluid_view(), view(bh, B), nff.p, Kff, cff.p,
                                             nfb.p, Kfb, cfb.p, nstat.p, d_ctr, stream);
            });
            if (coupled()) {
                nfr.reserve(size_t(Kfr) * std::max<uint32_t>(N, 1));
                cfr.reserve(std::max<uint32_t>(N, 1));
                T_(K_NEI_FLUID, [&] {
                    mps::launch::neighbors_ref(pos.p, scn(), uint32_t(cells()), g, fluid_view(), nfr.p, Kfr, cfr.p, nstat.p, d_ctr,
                                               stream);
                });
            }
            HIPCHK(hipMemcpyAsync(h_ctr, d_ctr, sizeof(NbrCounters), hipMemcpyDeviceToHost, stream));
            sync();
            bool grow = false;
            if (coupled() && h_ctr->max_fr > Kfr) { Kfr = (h_ctr->max_fr + 15u) & ~15u; grow = true; }
            if (h_ctr->max_ff > Kff) { Kff = (h_ctr->max_ff + 15u) & ~15u; grow = true; }
            if (h_ctr->max_fb > Kfb) { Kfb = (h_ctr->max_fb + 15u) & ~15u; grow = true; }
            if (do_b && h_ctr->max_bb > Kbb) { Kbb = (h_ctr->max_bb + 15u) & ~15u; grow = true; }
            if (!grow) {
                if (do_b) {
                    nnz_bb = h_ctr->nnz_bb;
                    if (nbobj > 1) {
                        bS4.reserve(std::max<uint32_t>(B, 1));
                        T_(K_BSUM, [&] {
                            mps::launch::boundary_sum_obj(cubic(), bpos.p, B, nbb.p, Kbb, cbb.p, bobj.p, bS4.p, stream);
                        });
                    } else {
                        bS.reserve(std::max<uint32_t>(B, 1));
                        T_(K_BSUM, [&] { mps::launch::boundary_sum(cubic(), bpos.p, B, nbb.p, Kbb, cbb.p, bS.p, stream); });
                    }
                    bnd_dirty = false;
                }
                nnz_ff = h_ctr->nnz_ff;
                nnz_fb = h_ctr->nnz_fb;
                pairs_valid = false;
                halo_valid = false;
                return;
            }
            nei_regrow++;
        }
        throw ApiError{MPS_ERR_STATE, "neighbour capacity did not converge"};
    }
    unsigned long long nnz_ff = 0, nnz_fb = 0, nnz_bb = 0;

    void boundary_volume() {
        if (nbobj > 1)
            T_(K_BVOL, [&] { mps::launch::boundary_volume_obj(bS4.p, bobj.p, nbobj, bvol.p, B, cfg.faithful_quirks, stream); });
        else
            T_(K_BVOL, [&] { mps::launch::boundary_volume(bS.p, bvol.p, B, cfg.faithful_quirks, stream); });
    }
    // The walls as k <= 4 objects of the given sizes (contiguous ranges in the current order).
    void set_boundary_objects(uint32_t k, const uint32_t* counts) {
        require_single("mps_set_boundary_objects");
        if (k == 0 || k > mps::kMaxWallObjects || !counts) throw ApiError{MPS_ERR_INVALID, "1 to 4 wall objects"};
        uint64_t total = 0;
        for (uint32_t o = 0; o < k; o++) total += counts[o];
        if (total != B) throw ApiError{MPS_ERR_INVALID, "object sizes do not add up to the wall particles"};
        if (morton_bits() + bits_for(k) > 31) throw ApiError{MPS_ERR_INVALID, "too many wall objects for the hash size"};
        std::vector<uint32_t> h(std::max<uint32_t>(B, 1), 0u);
        size_t at = 0;
        for (uint32_t o = 0; o < k; o++)
            for (uint32_t q = 0; q < counts[o]; q++) h[at++] = o;
        drop_rollout();
        bobj.reserve(h.size());
        up(bobj.p, h);
        nbobj = k;
        mesh.ok = false;
        bh.cells_init = false;
        bh.n_valid = 0;
        bnd_dirty = true;
        params_dirty = true;
    }
    void density_alpha(int flags) {
        ensure_halo();
        if (flags & (mps::launch::kDoAlpha | mps::launch::kDoPairs)) {
            flags |= mps::launch::kDoPairs;
            const size_t m = std::max<uint32_t>(N, 1);
            pairG_ff.reserve(size_t(Kff) * m);
            pairG_fb.reserve(size_t(Kfb) * m);
        }
        T_(K_DENSITY, [&] {
            mps::launch::density_alpha(cubic(), flags, scn(), nbr(), halo(), pos.p, vel.p, bpos.p, bvol.p, rho.p, alpha.p, color.p,
                                       pairG_ff.p, pairG_fb.p, stream);
        });
        if (flags & mps::launch::kDoPairs) pairs_valid = true;
    }
    void ensure_pairs() {
        if (!pairs_valid) density_alpha(mps::launch::kDoPairs);
    }
    // Block halos of the fluid lists for the halo-staged kernels (pressure solves, A.p)
    mps::Halo halo() const { return mps::Halo{halo_idx.p, halo_n.p, nb16.p}; }
    void ensure_halo() {
        if (halo_valid) return;
        const size_t m = std::max<uint32_t>(N, 1);
        const uint32_t nblk = uint32_t((size_t(N) + mps::dev::kBlock - 1) / mps::dev::kBlock);
        nb16.reserve(size_t(Kff) * m);
        halo_idx.reserve(size_t(std::max<uint32_t>(nblk, 1)) * mps::kHaloCap);
        halo_n.reserve(std::max<uint32_t>(nblk, 1));
        T_(K_HALO, [&] {
            mps::launch::halo_build(scn(), nbr(), nblk, cfg.halo_limit, halo_idx.p, halo_n.p, nb16.p, d_flags + 4, stream);
        });
        halo_valid = true;
    }

    // number of scenes whose loop `idx` is still running
    uint32_t fetch_active(int idx) {
        HIPCHK(hipMemcpyAsync(&h_active[idx], &d_active[idx], sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
        sync();
        return h_active[idx];
    }
    // Copy the S records of loop `idx` back and return the reference loop variable of each
    // scene (cap + 1 when it hit the cap), the largest one in *lmax.
    std::vector<uint32_t> fetch_records(int idx, uint32_t cap, uint32_t* lmax) {
        SolverDev* h = h_solver + size_t(idx) * S;
        HIPCHK(hipMemcpyAsync(h, d_solver + size_t(idx) * S, S * sizeof(SolverDev), hipMemcpyDeviceToHost, stream));
        sync();
        std::vector<uint32_t> l(S);
        *lmax = 0;
        for (uint32_t q = 0; q < S; q++) {
            l[q] = h[q].done ? h[q].l : cap + 1;
            *lmax = std::max(*lmax, l[q]);
        }
        return l;
    }
    // DF / CD loop (MPSSPHDFSPHUtil.cu:270-441)
    uint32_t pressure_solve(bool df, mps_step_stats* st) {
        const int idx = df ? 0 : 1;
        if (N == 0) return 0;
        SolverDev* sd = d_solver + size_t(idx) * S;
        mps::launch::solver_reset(sd, S, d_active + idx, stream);
        ensure_partials();
        ensure_pairs();
        ensure_halo();
        const mps::Scenes P = scn();
        const mps::Halo hl = halo();
        const double tol = df ? cfg.df_tol : cfg.cd_tol;
        const uint32_t lmin = df ? cfg.df_min_iter : cfg.cd_min_iter;
        const uint32_t cap = cfg.max_iter;
        uint32_t it = 1;
        bool done = false;
        while (it <= cap && !done) {
            const uint32_t chunk = chunk_len(it, std::max<uint32_t>(pred[idx], lmin + 1), cap - it + 1);
            std::vector<uint64_t> key = key_base(df ? 1 : 2, chunk, P, nbr());
            for (const void* q : {(const void*)pos.p, (const void*)vel.p, (const void*)bpos.p, (const void*)bvol.p,
                                  (const void*)bvel.p, (const void*)rho.p, (const void*)alpha.p, (const void*)pairG_ff.p,
                                  (const void*)pairG_fb.p, (const void*)kappa.p, (const void*)delta.p, (const void*)sd,
                                  (const void*)partials.p, (const void*)hl.idx, (const void*)hl.n, (const void*)hl.nb16})
                key.push_back(u(q));
            key.push_back(u(tol));
            key.push_back(lmin);
            launch_chunk(std::move(key), [&] {
                for (uint32_t c = 0; c < chunk; c++) {
                    T_(df ? K_EVAL_DF : K_EVAL_CD, [&] {
                        mps::launch::pressure_eval(cubic(), df, P, nbr(), hl, pos.p, vel.p, bpos.p, bvol.p, bvel.p, rho.p, alpha.p,
                                                   pairG_ff.p, pairG_fb.p, kappa.p, delta.p, sd, partials.p, stream);
                    }, false);
                    T_(K_FOLD, [&] {
                        mps::launch::fold_pressure(P, partials.p, sd, d_active + idx, tol, lmin, coupled(), stream);
                    }, false);
                    T_(K_APPLY, [&] {
                        mps::launch::pressure_apply(cubic(), P, nbr(), hl, pos.p, vel.p, bpos.p, bvol.p, pairG_ff.p, pairG_fb.p, kappa.p,
                                                    sd, stream);
                    }, false);
                    rec_snap();
                }
            });
            it += chunk;
            done = fetch_active(idx) == 0;
        }
        uint32_t l = 0;
        const std::vector<uint32_t> ls = fetch_records(idx, cap, &l);
        pred[idx] = std::max<uint32_t>(l, 1);
        if (rec) {
            if (df) {
                rec->ldf = ls;
                rec->df_last = rec->nsnap - 1;
            } else {
                rec->lcd = ls;
            }
        }
        if (prof_on) {  // eval and fold 1..min(l,cap), apply 1..l-1 did work (slowest scene)
            kuse[df ? K_EVAL_DF : K_EVAL_CD] += std::min(l, cap);
            kuse[K_FOLD] += std::min(l, cap);
            kuse[K_APPLY] += l - 1;
        }
        for (uint32_t q = 0; q < S && q < sstats.size(); q++) {
            mps_step_stats& o = sstats[q];
            const SolverDev& r = h_solver[size_t(idx) * S + q];
            double* log = df ? o.err_df : o.err_cd;
            const uint32_t nlog = std::min<uint32_t>(std::min(ls[q], cap), MPS_MAX_ITER_LOG);
            for (uint32_t k = 0; k < nlog; k++) log[k] = r.log[k];
            if (df) { o.l_df = ls[q]; o.hit_cap_df = ls[q] > cap; }
            else { o.l_cd = ls[q]; o.hit_cap_cd = ls[q] > cap; }
        }
        if (st) {
            if (df) { st->l_df = l; st->hit_cap_df = l > cap; }
            else { st->l_cd = l; st->hit_cap_cd = l > cap; }
            if (S == 1) {
                double* log = df ? st->err_df : st->err_cd;
                const double* src = df ? sstats[0].err_df : sstats[0].err_cd;
                std::memcpy(log, src, sizeof(double) * MPS_MAX_ITER_LOG);
            }
        }
        return l;
    }
    // ApplyImplicitViscosityNSurfaceTension + CGSolver (MPSSPHNonPressureUtil.cu:127-188,
    // MPSSolver.cuh:12-73)
    uint32_t viscosity_cg(mps_step_stats* st) {
        if (N == 0) return 0;
        SolverDev* sd = d_solver + 2 * size_t(S);
        uint32_t* act = d_active + 2;
        ensure_partials();
        ensure_pairs();
        const mps::Scenes P = scn();
        const size_t m = std::max<uint32_t>(N, 1);
        mps::launch::CgOperator op;
        op.recompute = cfg.cg_stream_coeffs ? 0 : 1;
        op.cubic = cubic();
        {  // halo depth of the streamed operator (the last list build's halos decide when asked to)
            const uint64_t nb = (uint64_t(N) + mps::dev::kBlock - 1) / mps::dev::kBlock;
            op.deep = !op.recompute && (cfg.cg_halo_depth == 2 || (cfg.cg_halo_depth == 0 && uint64_t(deep_halos) * 1000u > nb));
        }
        if (op.recompute) {
            pairA.reserve(size_t(Kff) * m);
            irho.reserve(m);
            op.ca = pairA.p;
            op.vel = vel.p;
            op.rho = rho.p;
            op.irho = irho.p;
        } else {
            pairAB.reserve(size_t(Kff) * m);
            op.cab = pairAB.p;
        }
        v3x.reserve(m); v3r.reserve(m); v3Ap.reserve(m); v3p.reserve(m); v3q.reserve(m);
        ensure_halo();
        const mps::Halo hl = halo();
        T_(K_ST_FACTOR, [&] { mps::launch::st_factor(P, nbr(), hl, pos.p, vel.p, pairG_ff.p, fst.p, cgx.p, stream); });
        T_(K_CG_SETUP, [&] {
            mps::launch::cg_setup(cubic(), P, nbr(), hl, pos.p, vel.p, rho.p, fst.p, op.recompute ? pairA.p : nullptr,
                                  op.recompute ? nullptr : pairAB.p, cgb.p, op.recompute ? irho.p : nullptr, stream);
        });
        mps::launch::solver_reset(sd, S, act, stream);
        const double tol = cfg.cg_tol;
        // p of iteration k is written to pbuf(k) while pbuf(k - 1) is read; p_0 = r
        auto pbuf = [&](uint32_t k) { return (k & 1u) ? v3q.p : v3p.p; };
        T_(K_CG_INIT, [&] {
            mps::launch::cg_apply(0, op, P, nbr(), hl, pos.p, cgx.p, nullptr, nullptr, pbuf(0), nullptr, cgb.p, v3r.p,
                                  v3x.p, sd, partials.p, stream);
        });
        T_(K_FOLD, [&] { mps::launch::fold_cg(0, P, partials.p, sd, act, tol, stream); });
        const uint32_t cap = cfg.cg_max_iter;
        uint32_t it = 1;
        bool done = false;
        // all iterations in one cooperative launch when every workgroup fits the device at once
        const uint32_t nblk = uint32_t((size_t(N) + mps::dev::kBlock - 1) / mps::dev::kBlock);
        const int coh = cfg.persistent_solver == 2 ? 1 : 0;
        const int variant = (op.recompute ? (op.cubic ? 1 : 2) : 0) + 3 * coh;
        if (cfg.persistent_solver && !prof_on && nblk > 0) {
            if (!persist_cap[variant]) persist_cap[variant] = std::max<uint32_t>(mps::launch::cg_persistent_capacity(op, coh), 1u);
            if (nblk <= persist_cap[variant] && nblk <= mps::kGridGroup * mps::kGridGroups) {
                mps::launch::zero_words(reinterpret_cast<uint32_t*>(d_grid), sizeof(mps::GridSync) / 4, stream);
                const uint32_t spin = cfg.persistent_spin ? cfg.persistent_spin : (1u << 22);
                mps::launch::cg_persistent(op, coh, P, nbr(), hl, pos.p, v3r.p, pbuf(0), pbuf(1), v3Ap.p, v3x.p, sd, act,
                                           partials.p, partials.p + nblk, d_grid, cap, tol, spin, stream);
                mps::GridSync g{};
                HIPCHK(hipMemcpyAsync(&g, d_grid, sizeof(g), hipMemcpyDeviceToHost, stream));
                sync();
                if (!g.gave_up) {
                    it = cap + 1;  // the launch ran every iteration the loop needed
                    if (st) st->cg_persistent = 1;
                } else {  // a barrier wait gave up: the whole solve again with chunked launches
                    persist_fallbacks++;
                    if (st) st->cg_persistent = 2;
                    mps::launch::solver_reset(sd, S, act, stream);
                    T_(K_CG_INIT, [&] {
                        mps::launch::cg_apply(0, op, P, nbr(), hl, pos.p, cgx.p, nullptr, nullptr, pbuf(0), nullptr, cgb.p,
                                              v3r.p, v3x.p, sd, partials.p, stream);
                    });
                    T_(K_FOLD, [&] { mps::launch::fold_cg(0, P, partials.p, sd, act, tol, stream); });
                }
            }
        }
        while (it <= cap && !done) {
            const uint32_t chunk = chunk_len(it, pred[2], cap - it + 1);
            std::vector<uint64_t> key = key_base(3, chunk, P, nbr());
            for (const void* q : {(const void*)op.cab, (const void*)op.ca, (const void*)op.vel, (const void*)op.rho,
                                  (const void*)op.irho, (const void*)pos.p, (const void*)pbuf(it - 1), (const void*)pbuf(it),
                                  (const void*)v3Ap.p, (const void*)v3r.p, (const void*)v3x.p, (const void*)sd,
                                  (const void*)partials.p, (const void*)act, (const void*)hl.idx, (const void*)hl.n,
                                  (const void*)hl.nb16})
                key.push_back(u(q));
            key.push_back(u(tol));
            // the allocation-free step keeps the fold launches (their counters would be scratch)
            const bool fold_in = cfg.cg_fold_in_kernel != 0 && !g_no_alloc;
            uint32_t* arr = fold_in ? arrive_counters() : nullptr;
            key.push_back(fold_in);
            key.push_back(u(arr));
            key.push_back(uint64_t(op.deep));
            launch_chunk(std::move(key), [&] {
                for (uint32_t c = 0; c < chunk; c++) {
                    const uint32_t k = it + c;
                    if (fold_in) {  // each launch's last workgroup folds (two launches per iteration)
                        T_(K_CG_APPLY, [&] {
                            mps::launch::cg_apply_fold(op, P, nbr(), hl, pos.p, v3r.p, pbuf(k - 1), pbuf(k), v3Ap.p, sd,
                                                       partials.p, arr, stream);
                        }, false);
                        T_(K_CG_UPDATE, [&] {
                            mps::launch::cg_update3_fold(P, v3x.p, v3r.p, pbuf(k), v3Ap.p, sd, partials.p, act, tol, arr,
                                                         stream);
                        }, false);
                        continue;
                    }
                    T_(K_CG_APPLY, [&] {
                        mps::launch::cg_apply(1, op, P, nbr(), hl, pos.p, nullptr, v3r.p, pbuf(k - 1), pbuf(k), v3Ap.p,
                                              nullptr, nullptr, nullptr, sd, partials.p, stream);
                    }, false);
                    T_(K_FOLD, [&] { mps::launch::fold_cg(1, P, partials.p, sd, act, tol, stream); }, false);
                    T_(K_CG_UPDATE, [&] {
                        mps::launch::cg_update3(P, v3x.p, v3r.p, pbuf(k), v3Ap.p, sd, partials.p, stream);
                    }, false);
                    T_(K_FOLD, [&] { mps::launch::fold_cg(2, P, partials.p, sd, act, tol, stream); }, false);
                }
            });
            it += chunk;
            done = fetch_active(2) == 0;
        }
        uint32_t l = 0;
        const std::vector<uint32_t> ls = fetch_records(2, cap, &l);
        pred[2] = std::max<uint32_t>(l, 1);
        if (prof_on) {  // A.p (with the direction) and the update ran for iterations 1..l-1
            kuse[K_CG_APPLY] += std::min(l - 1, cap);
            kuse[K_CG_UPDATE] += std::min(l - 1, cap);
            if (!cfg.cg_fold_in_kernel || g_no_alloc) kuse[K_FOLD] += 2 * std::min(l - 1, cap);
        }
        T_(K_COPY_OUT, [&] { mps::launch::copy_vel_out3(P, v3x.p, vel.p, stream); });
        for (uint32_t q = 0; q < S && q < sstats.size(); q++) {
            mps_step_stats& o = sstats[q];
            const SolverDev& r = h_solver[2 * size_t(S) + q];
            const uint32_t nlog = std::min<uint32_t>(std::min(ls[q], cap), MPS_MAX_ITER_LOG);
            for (uint32_t k = 0; k < nlog; k++) o.rr_cg[k] = r.log[k];
            o.l_cg = ls[q];
            o.hit_cap_cg = ls[q] > cap;
        }
        if (st) {
            if (S == 1) std::memcpy(st->rr_cg, sstats[0].rr_cg, sizeof(double) * MPS_MAX_ITER_LOG);
            st->l_cg = l;
            st->hit_cap_cg = l > cap;
        }
        return l;
    }
    // ApplyImplicitJacobiViscosity (SPHUtil.cu:344-453, commented out in the reference)
    uint32_t viscosity_jacobi(mps_step_stats* st) {
        if (N == 0) return 0;
        if (S > 1) throw ApiError{MPS_ERR_STATE, "the Jacobi viscosity solve runs one scene per context"};
        ensure_partials();
        const mps::Scenes P = scn();
        jprev.reserve(N); jcand.reserve(N);
        for (auto& t : jt) t.reserve(N);
        HIPCHK(hipMemcpyAsync(cgx.p, vel.p, sizeof(float4) * N, hipMemcpyDeviceToDevice, stream));
        HIPCHK(hipMemcpyAsync(jprev.p, vel.p, sizeof(float4) * N, hipMemcpyDeviceToDevice, stream));
        const double underRelax = 0.9, delta_ = 0.005;
        double omega = 1.0, rhoS = 0.995;
        double om[3] = {1.0, 1.0, 1.0};
        const uint32_t cap = cfg.jacobi_max_iter;
        uint32_t l = 0;
        bool conv = false;
        while (l < cap) {
            uint32_t minID = 0;
            double errs[3] = {0, 0, 0};
            if (l < 10u) {
                mps::launch::jacobi_candidate(cubic(), P, nbr(), pos.p, vel.p, rho.p, cgx.p, jt[0].p, stream);
                mps::launch::jacobi_error(cubic(), P, nbr(), pos.p, vel.p, rho.p, jt[0].p, jt[0].p, jt[0].p, 1, partials.p,
                                          d_err3, stream);
                HIPCHK(hipMemcpyAsync(h_err3, d_err3, 3 * sizeof(double), hipMemcpyDeviceToHost, stream));
                sync();
                errs[0] = h_err3[0];
            } else {
                if (l == 10u) {
                    om[0] = 2.0 / (2.0 - rhoS * rhoS);
                    om[1] = 2.0 / (2.0 - (rhoS + delta_) * (rhoS + delta_));
                    om[2] = 2.0 / (2.0 - (rhoS - delta_) * (rhoS - delta_));
                } else {
                    om[0] = 4.0 / (4.0 - rhoS * rhoS * omega);
                    om[1] = 4.0 / (4.0 - (rhoS + delta_) * (rhoS + delta_) * omega);
                    om[2] = 4.0 / (4.0 - (rhoS - delta_) * (rhoS - delta_) * omega);
                }
                const float w3[3] = {float(om[0]), float(om[1]), float(om[2])};
                mps::launch::jacobi_candidate(cubic(), P, nbr(), pos.p, vel.p, rho.p, cgx.p, jcand.p, stream);
                mps::launch::jacobi_relax(N, jcand.p, cgx.p, jprev.p, float(underRelax), w3, jt[0].p, jt[1].p, jt[2].p, stream);
                mps::launch::jacobi_error(cubic(), P, nbr(), pos.p, vel.p, rho.p, jt[0].p, jt[1].p, jt[2].p, 3, partials.p,
                                          d_err3, stream);
                HIPCHK(hipMemcpyAsync(h_err3, d_err3, 3 * sizeof(double), hipMemcpyDeviceToHost, stream));
                sync();
                for (int k = 0; k < 3; k++) errs[k] = h_err3[k];
                if (errs[minID] > errs[1]) minID = 1;
                if (errs[minID] > errs[2]) minID = 2;
                if (minID == 1) rhoS += delta_;
                else if (minID == 2) rhoS -= delta_;
                rhoS = std::min(std::max(0.5 + delta_, rhoS), 1.0 - delta_);
                omega = double(float(om[minID]));
            }
            // ApplyJacobiViscosity_kernel: previous <- current, current <- chosen candidate
            std::swap(jprev.p, cgx.p); std::swap(jprev.cap, cgx.cap);
            HIPCHK(hipMemcpyAsync(cgx.p, jt[minID].p, sizeof(float4) * N, hipMemcpyDeviceToDevice, stream));
            const double e = errs[minID] * (double(float(phys.dt)) / double(N));
            if (st && l < MPS_MAX_ITER_LOG) st->rr_cg[l] = e;
            if (e < cfg.jacobi_tol) { conv = true; break; }
            l++;
        }
        mps::launch::copy_vel_out(P, cgx.p, vel.p, stream);
        if (!sstats.empty()) {
            sstats[0].l_cg = l;
            sstats[0].hit_cap_cg = conv ? 0u : 1u;
        }
        if (st) {
            st->l_cg = l;
            st->hit_cap_cg = conv ? 0u : 1u;
        }
        return l;
    }
    void viscosity(mps_step_stats* st) {
        if (cfg.visc_solver == MPS_VISC_CG) viscosity_cg(st);
        else if (cfg.visc_solver == MPS_VISC_JACOBI) viscosity_jacobi(st);
    }
    void advect() {
        T_(K_ADVECT, [&] { mps::launch::advect(scn(), pos.p, vel.p, d_flags, stream); });
    }

    template <class T>
    void dcopy(DBuf<T>& dst, const T* from, size_t n) {
        dst.reserve(std::max<size_t>(n, 1));
        if (n) HIPCHK(hipMemcpyAsync(dst.p, from, n * sizeof(T), hipMemcpyDeviceToDevice, stream));
    }
    void rec_snap() {
        if (!rec) return;
        if (rec->snaps.size() <= rec->nsnap) rec->snaps.emplace_back();
        dcopy(rec->snaps[rec->nsnap++], vel.p, N);
    }

    // ------------------------------------------------------------------ adjoint
    void save_ck(Checkpoint& c) {
        dcopy(c.pos, pos.p, N); dcopy(c.vel, vel.p, N); dcopy(c.perm, perm.p, N);
        dcopy(c.bpos, bpos.p, B); dcopy(c.bvel, bvel.p, B); dcopy(c.bcolor, bcolor.p, B); dcopy(c.bvol, bvol.p, B);
        dcopy(c.bface, bface.p, B); dcopy(c.bbcc, bbcc.p, B);
        c.frame = frame;
    }
    void restore_ck(const Checkpoint& c) {
        auto back = [&](auto& dst, const auto& from, size_t n) {
            if (n) HIPCHK(hipMemcpyAsync(dst.p, from.p, n * sizeof(*dst.p), hipMemcpyDeviceToDevice, stream));
        };
        back(pos, c.pos, N); back(vel, c.vel, N); back(perm, c.perm, N);
        back(bpos, c.bpos, B); back(bvel, c.bvel, B); back(bcolor, c.bcolor, B); back(bvol, c.bvol, B);
        back(bface, c.bface, B); back(bbcc, c.bbcc, B);
        mps::launch::invert(perm.p, N, inv.p, stream);
        frame = c.frame;
        pairs_valid = false;
        halo_valid = false;
        bnd_dirty = true;
    }
    // Forward n steps from the current state, keeping a checkpoint every `every` steps; the
    // positions (and velocities) after steps obs[0..n_obs) go to pos_out (vel_out), n_obs blocks
    // of Nref x 3 in reference order.
    void adj_rollout(uint32_t n, uint32_t every, uint32_t* counts, uint32_t n_obs = 0, const uint32_t* obs = nullptr,
                     float* pos_out = nullptr, float* vel_out = nullptr) {
        for (uint32_t k = 0; k < n_obs; k++)
            if (!obs || obs[k] < 1 || obs[k] > n) throw ApiError{MPS_ERR_INVALID, "observed step outside the rollout"};
        if (n_obs && !pos_out) throw ApiError{MPS_ERR_INVALID, "no output for the observed positions"};
        if (coupled()) throw ApiError{MPS_ERR_STATE, "the adjoint does not cover coupled scenes"};
        if (nbobj > 1) throw ApiError{MPS_ERR_STATE, "the adjoint does not cover several wall objects"};
        if (cfg.visc_solver != MPS_VISC_CG) throw ApiError{MPS_ERR_STATE, "the adjoint needs the CG viscosity solve"};
        if (N == 0 || n == 0) throw ApiError{MPS_ERR_INVALID, "nothing to roll out"};
        sync_params();
        {  // equal fluid radii within every scene: then G_ij = G_ji and the reverse kernels take the fast path
            std::vector<float4> hp = down(pos.p, N);
            adj_eq = true;
            for (const SceneHost& h : scene_list()) {
                const float r0 = hp[h.begin].w;
                for (uint32_t q = h.begin; q < h.begin + h.nf; q++) adj_eq = adj_eq && hp[q].w == r0;
            }
        }
        every = std::max<uint32_t>(every, 1);
        dcopy(perm0, perm.p, N);
        const uint32_t nseg = (n + every - 1) / every;
        if (ckpts.size() < nseg) ckpts.resize(nseg);
        mps_step_stats st;
        for (uint32_t t = 0; t < n; t++) {
            if (t % every == 0) save_ck(ckpts[t / every]);
            step(&st);
            if (st.nonfinite) throw ApiError{MPS_ERR_NONFINITE, "non-finite state during the rollout"};
            for (uint32_t k = 0; k < n_obs; k++) {
                if (obs[k] != t + 1) continue;
                size_t cnt = 0, eb = 0;
                const size_t len = 3 * size_t(Nref);
                array_io(MPS_ARR_FLUID_POS, &cnt, &eb, pos_out + k * len, nullptr, len * sizeof(float));
                if (vel_out) array_io(MPS_ARR_FLUID_VEL, &cnt, &eb, vel_out + k * len, nullptr, len * sizeof(float));
            }
            if (counts)
                for (uint32_t q = 0; q < S; q++) {
                    uint32_t* o = counts + 3 * (size_t(t) * S + q);
                    o[0] = sstats[q].l_df;
                    o[1] = sstats[q].l_cg;
                    o[2] = sstats[q].l_cd;
                }
        }
        save_ck(ck_end);
        sync();
        adj_n = n;
        adj_every = every;
        adj_N = N;
        adj_Nref = Nref;
        adj_B = B;
        adj_S = S;
    }
    // Transposed solve A_f^T lambda = adv (lambda in cgx), same chunked loop as the forward CG,
    // started from the later step's solution when there is one.
    void adj_solveT(const mps::Nbr& nb, const float4* p0) {
        const mps::Scenes P = scn();
        SolverDev* sd = d_solver + 2 * size_t(S);
        uint32_t* act = d_active + 2;
        const double tol = cfg.adj_cg_tol != 0.0 ? cfg.adj_cg_tol : -1.0e-12;
        mps::launch::solver_reset(sd, S, act, stream);
        if (lam_valid) {
            cgrecA.reserve(2 * std::max<uint32_t>(N, 1));
            mps::launch::cg_pack(N, p0, lam.p, cgrecA.p, stream);
            mps::launch::adj_cg_init_warm(P, nb, pairAB.p, cabT.p, cgrecA.p, adv.p, cgx.p, cgr.p, cgrecB.p, partials.p,
                                          stream);
            mps::launch::adj_fold_start(P, partials.p, sd, act, tol, stream);
        } else {
            mps::launch::adj_cg_init(P, p0, adv.p, cgx.p, cgr.p, cgrecB.p, partials.p, stream);
            mps::launch::fold_cg(0, P, partials.p, sd, act, tol, stream);
        }
        const uint32_t cap = cfg.adj_cg_max_iter ? cfg.adj_cg_max_iter : 300u;
        uint32_t it = 1;
        bool done = fetch_active(2) == 0;
        while (it <= cap && !done) {
            const uint32_t chunk = chunk_len(it, pred_adj, cap - it + 1);
            std::vector<uint64_t> key = key_base(4, chunk, P, nb);
            for (const void* q : {(const void*)pairAB.p, (const void*)cabT.p, (const void*)cgrecB.p, (const void*)cgAp.p,
                                  (const void*)cgx.p, (const void*)cgr.p, (const void*)sd, (const void*)partials.p,
                                  (const void*)act})
                key.push_back(u(q));
            key.push_back(u(tol));
            launch_chunk(std::move(key), [&] {
                for (uint32_t c = 0; c < chunk; c++) {
                    mps::launch::adj_cg_applyT(P, nb, pairAB.p, cabT.p, cgrecB.p, cgAp.p, sd, partials.p, stream);
                    mps::launch::fold_cg(1, P, partials.p, sd, act, tol, stream);
                    mps::launch::cg_update(P, cgx.p, cgr.p, cgrecB.p, cgAp.p, sd, partials.p, stream);
                    mps::launch::fold_cg(2, P, partials.p, sd, act, tol, stream);
                    mps::launch::cg_direction(P, cgr.p, cgrecB.p, sd, stream);
                }
            });
            it += chunk;
            done = fetch_active(2) == 0;
        }
        uint32_t l = 0;
        (void)fetch_records(2, cap, &l);
        pred_adj = std::max<uint32_t>(l, 1);
    }
    // Reverse of one recorded step: (adx, adv) from the step's end state to its start, in the
    // slot order of the previous step.
    void reverse_step(TapeStep& T) {
        const mps::Scenes P = scn();
        const mps::Nbr nb{T.nff.p, T.cff.p, T.nfb.p, T.cfb.p};
        const size_t m = std::max<uint32_t>(N, 1);
        pairG_ff.reserve(size_t(T.Kff) * m);
        pairG_fb.reserve(size_t(T.Kfb) * m);
        pairs_valid = false;
        halo_valid = false;
        const mps::launch::AdjKind ak{cubic(), adj_eq ? 1 : 0};
        mps::launch::density_alpha(cubic(), mps::launch::kDoDensity | mps::launch::kDoAlpha | mps::launch::kDoPairs, P, nb,
                                   mps::Halo{}, T.pos.p, T.snaps[0].p, T.bpos.p, T.bvol.p, rho.p, alpha.p, color.p, pairG_ff.p,
                                   pairG_fb.p, stream);
        abar.reserve(m); rbar.reserve(m); dbar.reserve(m); fbar.reserve(m); fstr.reserve(m);
        aux4.reserve(m); gvec.reserve(m); avin.reserve(m); adx2.reserve(m); adv2.reserve(m);
        mps::launch::zero_words(reinterpret_cast<uint32_t*>(abar.p), N, stream);
        mps::launch::zero_words(reinterpret_cast<uint32_t*>(rbar.p), N, stream);
        mps::launch::adj_advect(P, T.pos.p, T.snaps[T.nsnap - 1].p, adx.p, adv.p, stream);
        auto loop_rev = [&](int df, const std::vector<uint32_t>& L, uint32_t base) {
            uint32_t lmax = 0;
            for (uint32_t l : L) lmax = std::max(lmax, l);
            if (lmax < 2) return;
            d_L.reserve(S);
            up(d_L.p, L);
            for (uint32_t it = lmax - 1; it >= 1; it--) {
                const float4* vk = T.snaps[base + it - 1].p;
                mps::launch::adj_peval(df, P, nb, T.pos.p, vk, T.bpos.p, T.bvol.p, bvel.p, rho.p, alpha.p, pairG_ff.p,
                                       pairG_fb.p, aux4.p, stream);
                mps::launch::adj_kbar(ak, df, P, nb, T.pos.p, T.bpos.p, T.bvol.p, pairG_ff.p, pairG_fb.p, aux4.p, alpha.p, adv.p,
                                      d_L.p, it, dbar.p, adx.p, abar.p, rbar.p, stream);
                mps::launch::adj_dbar(ak, P, nb, T.pos.p, vk, T.bpos.p, T.bvol.p, bvel.p, pairG_ff.p, pairG_fb.p, dbar.p, d_L.p, it,
                                      adv.p, adx.p, stream);
            }
        };
        loop_rev(0, T.lcd, T.cg_out);
        {  // the implicit viscosity / surface-tension solve
            const float4* vin = T.snaps[T.df_last].p;
            const float4* vst = T.snaps[T.cg_out].p;
            pairAB.reserve(size_t(T.Kff) * m); cabT.reserve(size_t(T.Kff) * m); rev.reserve(size_t(T.Kff) * m);
            cgrecB.reserve(2 * m);
            mps::launch::st_factor(P, nb, mps::Halo{}, T.pos.p, vin, pairG_ff.p, fstr.p, cgx.p, stream);
            mps::launch::cg_setup(cubic(), P, nb, mps::Halo{}, T.pos.p, vin, rho.p, fstr.p, nullptr, pairAB.p, cgb.p, nullptr,
                                  stream);
            mps::launch::zero_words(d_flags + 2, 1, stream);
            mps::launch::adj_rev_index(P, nb, rev.p, d_flags + 2, stream);
            mps::launch::adj_transpose_coeffs(P, nb, rev.p, pairAB.p, cabT.p, stream);
            adj_solveT(nb, T.pos.p);
            lam.reserve(m);
            HIPCHK(hipMemcpyAsync(lam.p, cgx.p, sizeof(float4) * N, hipMemcpyDeviceToDevice, stream));
            lam_valid = true;
            mps::launch::adj_cg_grad(ak, P, nb, T.pos.p, vin, vst, rho.p, fstr.p, cgx.p, avin.p, adx.p, rbar.p, fbar.p, partials.p,
                                     stream);
            mps::launch::adj_fold2(P, partials.p, acc_mu.p, acc_sig.p, stream);
            mps::launch::adj_fst(ak, P, nb, T.pos.p, vin, pairG_ff.p, fbar.p, avin.p, adx.p, stream);
            std::swap(adv, avin);
        }
        loop_rev(1, T.ldf, 0);
        mps::launch::adj_gvec(P, nb, T.pos.p, T.bpos.p, T.bvol.p, pairG_ff.p, pairG_fb.p, abar.p, gvec.p, stream);
        mps::launch::adj_density(ak, P, nb, T.pos.p, T.bpos.p, T.bvol.p, pairG_ff.p, pairG_fb.p, rbar.p, gvec.p, adx.p, stream);
        lam2.reserve(m);
        mps::launch::adj_unpermute(N, T.src.p, adx.p, adv.p, lam_valid ? lam.p : nullptr, adx2.p, adv2.p, lam2.p, stream);
        std::swap(adx, adx2);
        std::swap(adv, adv2);
        std::swap(lam, lam2);
    }
    // Gradients of a loss of the states observed after steps obs[0..n_obs) of the rollout
    // (cotangents gx, gv: n_obs blocks of Nref x 3 in the reference order of each observed state;
    // gv may be null) with respect to the velocities at its start and each scene's viscosity and
    // surface-tension coefficient. Each observation's cotangent is added to the adjoint state
    // just before the step that produced that state is reversed.
    void adj_backward(uint32_t n_obs, const uint32_t* obs, const float* gx, const float* gv, float* gv0, double* gmu,
                      double* gsig) {
        if (!adj_n) throw ApiError{MPS_ERR_STATE, "no rollout to differentiate"};
        if (N != adj_N || Nref != adj_Nref || B != adj_B || S != adj_S || ckpts.size() * adj_every < adj_n)
            throw ApiError{MPS_ERR_STATE, "the particles changed since the rollout"};
        if (!gx || !gv0 || n_obs == 0 || !obs) throw ApiError{MPS_ERR_INVALID, "no cotangents"};
        std::vector<std::vector<uint32_t>> at(adj_n + 1);  // observation indices per observed step
        for (uint32_t k = 0; k < n_obs; k++) {
            if (obs[k] < 1 || obs[k] > adj_n) throw ApiError{MPS_ERR_INVALID, "observed step outside the rollout"};
            at[obs[k]].push_back(k);
        }
        sync_params();
        ensure_partials();
        const size_t m = std::max<uint32_t>(N, 1);
        adx.reserve(m); adv.reserve(m);
        mps::launch::zero_words(reinterpret_cast<uint32_t*>(adx.p), 4 * size_t(N), stream);
        mps::launch::zero_words(reinterpret_cast<uint32_t*>(adv.p), 4 * size_t(N), stream);
        const size_t obs_len = 3 * size_t(Nref);
        gobs.reserve(2 * std::max<size_t>(obs_len, 1));
        lam_valid = false;
        acc_mu.reserve(S); acc_sig.reserve(S);
        mps::launch::zero_words(reinterpret_cast<uint32_t*>(acc_mu.p), 2 * size_t(S), stream);
        mps::launch::zero_words(reinterpret_cast<uint32_t*>(acc_sig.p), 2 * size_t(S), stream);
        const uint32_t nseg = (adj_n + adj_every - 1) / adj_every;
        mps_step_stats st;
        for (uint32_t sg = nseg; sg-- > 0;) {
            restore_ck(ckpts[sg]);
            const uint32_t t0 = sg * adj_every, t1 = std::min(adj_n, t0 + adj_every);
            if (tape.size() < t1 - t0) tape.resize(t1 - t0);
            for (uint32_t t = t0; t < t1; t++) {
                rec = &tape[t - t0];
                rec->nsnap = 0;
                try {
                    step(&st);
                } catch (...) {
                    rec = nullptr;
                    throw;
                }
                rec = nullptr;
            }
            for (uint32_t t = t1; t-- > t0;) {
                TapeStep& T = tape[t - t0];
                for (uint32_t k : at[t + 1]) {  // the state after step t, in that step's slot order
                    HIPCHK(hipMemcpyAsync(gobs.p, gx + k * obs_len, obs_len * sizeof(float), hipMemcpyHostToDevice, stream));
                    if (gv)
                        HIPCHK(hipMemcpyAsync(gobs.p + obs_len, gv + k * obs_len, obs_len * sizeof(float), hipMemcpyHostToDevice,
                                              stream));
                    mps::launch::adj_add_cotangent(N, T.perm_end.p, gobs.p, gv ? gobs.p + obs_len : nullptr, adx.p, adv.p, stream);
                    sync();  // the host buffers may be reused by the next observation's copy
                }
                reverse_step(T);
            }
        }
        HIPCHK(hipMemcpyAsync(h_flags + 2, d_flags + 2, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
        sync();
        if (h_flags[2]) throw ApiError{MPS_ERR_STATE, "asymmetric neighbour lists in the adjoint"};
        std::vector<float4> hv = down(adv.p, N);
        std::vector<uint32_t> hp0 = down(perm0.p, N);
        for (uint32_t q = 0; q < N; q++) {
            const uint32_t r = hp0[q];
            if (r == mps::dev::kEmpty) continue;
            gv0[3 * size_t(r)] = hv[q].x;
            gv0[3 * size_t(r) + 1] = hv[q].y;
            gv0[3 * size_t(r) + 2] = hv[q].z;
        }
        std::vector<double> hm = down(acc_mu.p, S), hs = down(acc_sig.p, S);
        for (uint32_t q = 0; q < S; q++) {
            if (gmu) gmu[q] = hm[q];
            if (gsig) gsig[q] = hs[q];
        }
        restore_ck(ck_end);
        sync();
    }

    // ------------------------------------------------------------------ autotuning
    void autotune(uint32_t steps, mps_tune_result* out) {
        if (N == 0 || steps == 0) throw ApiError{MPS_ERR_INVALID, "nothing to tune"};
        sync_params();
        Checkpoint c;
        save_ck(c);
        mps_step_stats st;
        // ms per step of `steps` steps from the saved state under the current configuration
        auto time_steps = [&]() {
            restore_ck(c);
            step(&st);  // captures graphs, sizes buffers
            restore_ck(c);
            sync();
            const auto t0 = std::chrono::steady_clock::now();
            for (uint32_t q = 0; q < steps; q++) step(&st);
            sync();
            return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / steps;
        };
        mps_tune_result r{};
        // 1. graph replay and the host poll interval
        const uint32_t polls[4] = {0u, 1u, 4u, 16u};
        double best = 1e300;
        uint32_t bg = cfg.use_graphs, bh = cfg.host_sync_every, n = 0;
        for (uint32_t g = 0; g < 2; g++)
            for (uint32_t k = 0; k < 4; k++) {
                cfg.use_graphs = g;
                cfg.host_sync_every = polls[k];
                const double ms = time_steps();
                r.ms_per_step[n++] = ms;
                if (ms < best) {
                    best = ms;
                    bg = g;
                    bh = polls[k];
                }
            }
        cfg.use_graphs = bg;
        cfg.host_sync_every = bh;
        // 2. the viscosity operator's coefficient stream (only where the CG runs)
        uint32_t bc = cfg.cg_stream_coeffs;
        if (cfg.visc_solver == MPS_VISC_CG) {
            best = 1e300;
            for (uint32_t v = 0; v < 2; v++) {
                cfg.cg_stream_coeffs = 1u - v;
                const double ms = time_steps();
                r.ms_coeffs[v] = ms;
                n++;
                if (ms < best) {
                    best = ms;
                    bc = 1u - v;
                }
            }
            cfg.cg_stream_coeffs = bc;
        }
        // 3. the CG's folds: separate launches or in each launch's last workgroup
        uint32_t bf = cfg.cg_fold_in_kernel;
        if (cfg.visc_solver == MPS_VISC_CG) {
            best = 1e300;
            for (uint32_t v = 0; v < 2; v++) {
                cfg.cg_fold_in_kernel = v;
                const double ms = time_steps();
                r.ms_fold[v] = ms;
                n++;
                if (ms < best) {
                    best = ms;
                    bf = v;
                }
            }
            cfg.cg_fold_in_kernel = bf;
        }
        // 4. the persistent CG solver, where every workgroup fits the device at once
        uint32_t bp = 0;
        if (cfg.visc_solver == MPS_VISC_CG) {
            cfg.persistent_solver = 0;
            r.ms_persistent[0] = time_steps();
            n++;
            mps::launch::CgOperator op;
            op.recompute = cfg.cg_stream_coeffs ? 0 : 1;
            op.cubic = cubic();
            const int variant = (op.recompute ? (op.cubic ? 1 : 2) : 0) + 3;
            if (!persist_cap[variant]) persist_cap[variant] = std::max<uint32_t>(mps::launch::cg_persistent_capacity(op, 1), 1u);
            const uint32_t nblk = uint32_t((size_t(N) + mps::dev::kBlock - 1) / mps::dev::kBlock);
            if (nblk <= persist_cap[variant] && nblk <= mps::kGridGroup * mps::kGridGroups) {
                cfg.persistent_solver = 2;
                r.ms_persistent[1] = time_steps();
                n++;
                if (r.ms_persistent[1] < r.ms_persistent[0]) bp = 2;
            }
        }
        cfg.persistent_solver = bp;
        restore_ck(c);
        sync();
        c.release();
        if (out) {
            *out = r;
            out->use_graphs = bg;
            out->host_sync_every = bh;
            out->cg_stream_coeffs = bc;
            out->persistent_solver = bp;
            out->cg_fold_in_kernel = bf;
            out->n_tried = n;
        }
    }

    void begin_stats() {
        sstats.assign(S, mps_step_stats{});
        const std::vector<SceneHost> sl = scene_list();
        for (uint32_t q = 0; q < S; q++) {
            sstats[q].frame = frame;
            sstats[q].n_fluid = sl[q].nf;
            sstats[q].n_boundary = sl[q].nb;
        }
    }
    void run_phase(int32_t ph, mps_step_stats* st) {
        sync_params();
        if (sstats.size() != S) begin_stats();
        switch (ph) {
            case MPS_PHASE_ZSORT: zsort(); break;
            case MPS_PHASE_HASH: hash(); break;
            case MPS_PHASE_NEIGHBORS: neighbors(); break;
            case MPS_PHASE_BOUNDARY_VOLUME: boundary_volume(); break;
            case MPS_PHASE_DENSITY: density_alpha(mps::launch::kDoDensity); break;
            case MPS_PHASE_ALPHA: density_alpha(mps::launch::kDoAlpha); break;
            case MPS_PHASE_GRAVITY: density_alpha(mps::launch::kDoGravity); break;
            case MPS_PHASE_DIVERGENCE: pressure_solve(true, st); break;
            case MPS_PHASE_VISCOSITY: viscosity(st); break;
            case MPS_PHASE_CONSTANT_DENSITY: pressure_solve(false, st); break;
            case MPS_PHASE_ADVECT: advect(); break;
            default: throw ApiError{MPS_ERR_INVALID, "unknown phase"};
        }
    }
    void step(mps_step_stats* st) {
        mps_step_stats local;
        if (!st) st = &local;
        std::memset(st, 0, sizeof(*st));
        sync_params();
        begin_stats();
        st->frame = frame;
        st->n_fluid = Nref;
        st->n_boundary = B;
        const uint32_t regrow0 = nei_regrow;
        mps::launch::zero_words(d_flags, 1, stream);
        mps::launch::zero_words(d_flags + 4, 4, stream);
        if (cfg.zsort_interval && (frame % cfg.zsort_interval) == 1u) {
            zsort();
            st->zsorted = 1;
        }
        hash();
        neighbors();
        boundary_volume();
        if (rec) {
            rec->Kff = Kff;
            rec->Kfb = Kfb;
            dcopy(rec->pos, pos.p, N);
            dcopy(rec->nff, nff.p, size_t(Kff) * N);
            dcopy(rec->cff, cff.p, N);
            dcopy(rec->nfb, nfb.p, size_t(Kfb) * N);
            dcopy(rec->cfb, cfb.p, N);
            dcopy(rec->bpos, bpos.p, B);
            dcopy(rec->bvol, bvol.p, B);
        }
        density_alpha(mps::launch::kDoDensity | mps::launch::kDoAlpha | mps::launch::kDoGravity);
        rec_snap();
        pressure_solve(true, st);
        viscosity(st);
        if (rec) {
            rec_snap();
            rec->cg_out = rec->nsnap - 1;
        }
        pressure_solve(false, st);
        advect();
        if (rec) dcopy(rec->perm_end, perm.p, N);
        HIPCHK(hipMemcpyAsync(h_flags, d_flags, 8 * sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
        sync();
        st->nonfinite = h_flags[0] ? 1u : 0u;
        st->halo_max = h_flags[4];
        st->halo_spill = h_flags[5];
        st->halo_blocks = h_flags[6];
        deep_halos = h_flags[7];
        st->nnz_ff = nnz_ff;
        st->nnz_fb = nnz_fb;
        st->nnz_bb = nnz_bb;
        st->nei_regrow = nei_regrow - regrow0;
        for (mps_step_stats& o : sstats) {
            o.zsorted = st->zsorted;
            o.nonfinite = st->nonfinite;
            o.nei_regrow = st->nei_regrow;
            o.cg_persistent = st->cg_persistent;
        }
        if (S == 1) {  // totals of the one scene
            sstats[0].nnz_ff = nnz_ff;
            sstats[0].nnz_fb = nnz_fb;
            sstats[0].nnz_bb = nnz_bb;
        }
        frame++;
        sized_N = std::max(sized_N, N);
        sized_B = std::max(sized_B, B);
        sized_S = std::max(sized_S, S);
        if (st->zsorted) {
            zsorted_N = std::max(zsorted_N, N);
            zsorted_B = std::max(zsorted_B, B);
        }
        if (prof_on) prof_collect();
    }

    // One step that allocates nothing: the step's reduction scratch comes from the device's
    // `scratch` array, and any context buffer that would have to grow makes the call fail
    // (before any physics is applied) instead of allocating. Graph capture and kernel timing
    // allocate, so the step launches directly and timing must be off.
    void step_noalloc(mps_step_stats* st) {
        if (prof_on) throw ApiError{MPS_ERR_STATE, "mps_step_noalloc: kernel timing allocates events; disable it"};
        if (std::max(N, B) > kNoAllocSlots || S > kNoAllocScenes)
            throw ApiError{MPS_ERR_INVALID, "mps_step_noalloc: more than 8388608 particles per set or 64 scenes "
                                            "(the size of its scratch); use mps_step"};
        if (N > sized_N || B > sized_B || S > sized_S ||
            (cfg.zsort_interval && (frame % cfg.zsort_interval) == 1u && (N > zsorted_N || B > zsorted_B)))
            refuse_alloc();
        unsigned char* base = scratch_address(device);
        struct Swap {
            mps_ctx* c;
            double* partials;
            size_t partials_cap;
            uint2* nstat;
            size_t nstat_cap;
            NbrCounters* ctr;
            SolverDev* solver;
            uint32_t solver_cap;
            uint32_t* active;
            ~Swap() {
                g_no_alloc = false;
                if (!c) return;
                c->partials.p = partials; c->partials.cap = partials_cap;
                c->nstat.p = nstat; c->nstat.cap = nstat_cap;
                c->d_ctr = ctr; c->d_solver = solver; c->solver_cap = solver_cap; c->d_active = active;
            }
        } sw{nullptr, partials.p, partials.cap, nstat.p, nstat.cap, d_ctr, d_solver, solver_cap, d_active};
        g_no_alloc = true;
        sync_params();  // parameter upload; the solver records below replace the context's
        sw.c = this;
        partials.p = reinterpret_cast<double*>(base);
        partials.cap = kScrPartials;
        nstat.p = reinterpret_cast<uint2*>(base + kScrOffNstat);
        nstat.cap = kScrNstat;
        d_ctr = reinterpret_cast<NbrCounters*>(base + kScrOffCtr);
        d_solver = reinterpret_cast<SolverDev*>(base + kScrOffSolver);
        solver_cap = kNoAllocScenes;
        d_active = reinterpret_cast<uint32_t*>(base + kScrOffActive);
        // the records keep the last call's contents: start them as a fresh context would
        for (int k = 0; k < 3; k++) mps::launch::solver_reset(d_solver + size_t(k) * S, S, d_active + k, stream);
        step(st);
    }

    // ------------------------------------------------------------------ setup
    void alloc_fluid(uint32_t n) {
        const size_t m = std::max<uint32_t>(n, 1);
        pos.reserve(m); vel.reserve(m); color.reserve(m); cgb.reserve(m); cgx.reserve(m); cgr.reserve(m);
        cgAp.reserve(m); rho.reserve(m); alpha.reserve(m); kappa.reserve(m); delta.reserve(m); fst.reserve(m);
        cff.reserve(m); cfb.reserve(m); perm.reserve(m); inv.reserve(m); mkey.reserve(m);
    }
    void require_single(const char* what) const {
        if (!scenes.empty()) throw ApiError{MPS_ERR_STATE, std::string(what) + ": the context holds a batch of scenes"};
    }
    // Zero the derived fluid state of slots [0, n) and start at frame 0.
    void reset_fluid_state(uint32_t n) {
        if (n) {
            mps::launch::zero_words(reinterpret_cast<uint32_t*>(rho.p), n, stream);
            mps::launch::zero_words(reinterpret_cast<uint32_t*>(alpha.p), n, stream);
            mps::launch::zero_words(reinterpret_cast<uint32_t*>(kappa.p), n, stream);
            mps::launch::zero_words(reinterpret_cast<uint32_t*>(delta.p), n, stream);
            mps::launch::zero_words(reinterpret_cast<uint32_t*>(fst.p), n, stream);
            mps::launch::zero_words(reinterpret_cast<uint32_t*>(cgb.p), 4 * size_t(n), stream);
            mps::launch::zero_words(cff.p, n, stream);
            mps::launch::zero_words(cfb.p, n, stream);
            // Padding slots between scenes keep kEmpty in mkey and src for good: the hash
            // writes both only at live slots, and the reorder gathers skip kEmpty sources.
            mps::launch::fill_u32(mkey.p, n, mps::dev::kEmpty, stream);
            src.reserve(n);
            mps::launch::fill_u32(src.p, n, mps::dev::kEmpty, stream);
        }
        fh.cells_init = false;
        fh.n_valid = 0;
        pairs_valid = false;
        halo_valid = false;
        drop_rollout();
        frame = 0;
        pred[0] = 2; pred[1] = 3; pred[2] = 8; pred_adj = 8;
        params_dirty = true;
    }
    void alloc_boundary(uint32_t n) {
        drop_rollout();
        const size_t m = std::max<uint32_t>(n, 1);
        bpos.reserve(m); bvel.reserve(m); bcolor.reserve(m); bvol.reserve(m); bS.reserve(m); bface.reserve(m); bbcc.reserve(m);
        cbb.reserve(m);
    }
    void upload_fluid(size_t n, const float* p, const float* v, const float* r, const float* m) {
        if (n > 0x7FFFFFFFu) throw ApiError{MPS_ERR_INVALID, "too many particles"};
        require_single("mps_upload_fluid");
        alloc_fluid(uint32_t(n));
        std::vector<float4> hp(n), hv(n), hc(n, make_float4(0.f, 0.f, 1.f, 1.f));
        for (size_t i = 0; i < n; i++) {
            hp[i] = make_float4(p[3 * i], p[3 * i + 1], p[3 * i + 2], r[i]);
            hv[i] = make_float4(v ? v[3 * i] : 0.f, v ? v[3 * i + 1] : 0.f, v ? v[3 * i + 2] : 0.f, m ? m[i] : float(mMass));
        }
        N = Nref = uint32_t(n);
        if (n) {
            HIPCHK(hipMemcpyAsync(pos.p, hp.data(), n * sizeof(float4), hipMemcpyHostToDevice, stream));
            HIPCHK(hipMemcpyAsync(vel.p, hv.data(), n * sizeof(float4), hipMemcpyHostToDevice, stream));
            HIPCHK(hipMemcpyAsync(color.p, hc.data(), n * sizeof(float4), hipMemcpyHostToDevice, stream));
            mps::launch::iota(perm.p, uint32_t(n), stream);
            mps::launch::iota(inv.p, uint32_t(n), stream);
        }
        reset_fluid_state(uint32_t(n));
        sync();
    }

    // ------------------------------------------------------------------ batched scenes
    void batch_clear() {
        sync();
        drop_rollout();
        staged.clear();
        scenes.clear();
        sstats.clear();
        N = Nref = B = 0;
        S = 1;
        fh.cells_init = bh.cells_init = false;
        fh.n_valid = bh.n_valid = 0;
        bnd_dirty = true;
        pairs_valid = false;
        halo_valid = false;
        frame = 0;
        params_dirty = true;
    }
    void batch_add(const mps_physics& ph, const mps_sph_material& mt, size_t nf, const float* p, const float* v,
                   const float* r, const float* m, size_t nb, const float* bp, const float* br, const float* bm) {
        if (nf == 0) throw ApiError{MPS_ERR_INVALID, "a scene needs fluid particles"};
        if (!(ph.dt > 0.0) || !(mt.radius > 0.0) || !(mt.density > 0.0)) throw ApiError{MPS_ERR_INVALID, "bad physics or material"};
        StagedScene sc;
        sc.h.phys = ph;
        sc.h.mat = mt;
        sc.h.V = mt.radius * mt.radius * mt.radius / 48.0 * 3.141592;
        sc.h.mass = mt.density * sc.h.V;
        sc.h.nf = uint32_t(nf);
        sc.h.nb = uint32_t(nb);
        sc.pos.assign(p, p + 3 * nf);
        sc.rad.assign(r, r + nf);
        sc.vel = v ? std::vector<float>(v, v + 3 * nf) : std::vector<float>(3 * nf, 0.0f);
        sc.mass = m ? std::vector<float>(m, m + nf) : std::vector<float>(nf, float(sc.h.mass));
        if (nb) {
            sc.bpos.assign(bp, bp + 3 * nb);
            sc.brad.assign(br, br + nb);
            sc.bmass = bm ? std::vector<float>(bm, bm + nb) : std::vector<float>(nb, 0.0f);
        }
        staged.push_back(std::move(sc));
    }
    // Lay the staged scenes out back to back and upload them (replaces the current state).
    void batch_commit() {
        if (staged.empty()) throw ApiError{MPS_ERR_STATE, "no scenes were added"};
        const uint32_t blk = mps::dev::kBlock;
        uint64_t n = 0, nref = 0, nb = 0;
        std::vector<SceneHost> lay;
        for (StagedScene& sc : staged) {
            SceneHost h = sc.h;
            h.begin = uint32_t((n + blk - 1) / blk * blk);
            h.ref0 = uint32_t(nref);
            h.b0 = uint32_t(nb);
            n = uint64_t(h.begin) + h.nf;
            nref += h.nf;
            nb += h.nb;
            if (n > 0x7FFFFFFFu || nb > 0x7FFFFFFFu) throw ApiError{MPS_ERR_INVALID, "too many particles"};
            lay.push_back(h);
        }
        const uint32_t nsc = uint32_t(lay.size());
        if (cfg.couple_scenes)  // fluid objects of one world: one time step, gravity and box
            for (const SceneHost& h : lay)
                if (std::memcmp(&h.phys, &lay[0].phys, sizeof(mps_physics)) != 0)
                    throw ApiError{MPS_ERR_INVALID, "coupled scenes must share dt, gravity and the box"};
        if (morton_bits() + bits_for(nsc) > 31 || cell_bits() + bits_for(nsc) > 31 ||
            uint64_t(nsc) * cells() > 0xFFFFFFFFull)
            throw ApiError{MPS_ERR_INVALID, "too many scenes for the hash size"};
        N = uint32_t(n);
        Nref = uint32_t(nref);
        B = uint32_t(nb);
        mesh.ok = false;
        nbobj = 1;
        alloc_fluid(N);
        alloc_boundary(B);
        std::vector<float4> hp(N, make_float4(0.f, 0.f, 0.f, 0.f)), hv(N, make_float4(0.f, 0.f, 0.f, 0.f));
        std::vector<float4> hc(N, make_float4(0.f, 0.f, 1.f, 1.f));
        std::vector<uint32_t> hperm(N, mps::dev::kEmpty), hinv(std::max<uint32_t>(Nref, 1)), hbs(std::max<uint32_t>(B, 1), 0u);
        std::vector<float4> hbp(std::max<uint32_t>(B, 1)), hbv(std::max<uint32_t>(B, 1));
        const size_t nblk = std::max<size_t>((size_t(N) + blk - 1) / blk, 1);
        std::vector<uint32_t> hblk(nblk, 0u), hdelta(nsc);
        for (uint32_t q = 0; q < nsc; q++) {
            const SceneHost& h = lay[q];
            const StagedScene& sc = staged[q];
            for (uint32_t i = 0; i < h.nf; i++) {
                const uint32_t sl = h.begin + i;
                hp[sl] = make_float4(sc.pos[3 * i], sc.pos[3 * i + 1], sc.pos[3 * i + 2], sc.rad[i]);
                hv[sl] = make_float4(sc.vel[3 * i], sc.vel[3 * i + 1], sc.vel[3 * i + 2], sc.mass[i]);
                hperm[sl] = h.ref0 + i;
                hinv[h.ref0 + i] = sl;
            }
            for (uint32_t b = h.begin / blk; b < (h.begin + h.nf + blk - 1) / blk; b++) hblk[b] = q;
            hdelta[q] = h.begin - h.ref0;
            for (uint32_t i = 0; i < h.nb; i++) {
                hbp[h.b0 + i] = make_float4(sc.bpos[3 * i], sc.bpos[3 * i + 1], sc.bpos[3 * i + 2], sc.brad[i]);
                hbv[h.b0 + i] = make_float4(0.f, 0.f, 0.f, sc.bmass[i]);
                hbs[h.b0 + i] = q;
            }
        }
        blk_scene.reserve(nblk);
        d_delta.reserve(nsc);
        bscene.reserve(std::max<uint32_t>(B, 1));
        up(pos.p, hp); up(vel.p, hv); up(color.p, hc); up(perm.p, hperm); up(inv.p, hinv);
        up(blk_scene.p, hblk); up(d_delta.p, hdelta);
        if (B) {
            up(bpos.p, hbp); up(bvel.p, hbv); up(bscene.p, hbs);
            mps::launch::zero_words(reinterpret_cast<uint32_t*>(bcolor.p), 4 * size_t(B), stream);
            mps::launch::zero_words(reinterpret_cast<uint32_t*>(bvol.p), B, stream);
            mps::launch::zero_words(bface.p, B, stream);
            mps::launch::zero_words(reinterpret_cast<uint32_t*>(bbcc.p), 2 * size_t(B), stream);
            mps::launch::zero_words(cbb.p, B, stream);
        }
        scenes = lay;
        S = nsc;
        staged.clear();
        bh.cells_init = false;
        bh.n_valid = 0;
        bnd_dirty = true;
        reset_fluid_state(N);
        sstats.clear();
        sync();
    }
    void upload_boundary(size_t n, const float* p, const float* r, const float* m) {
        if (n > 0x7FFFFFFFu) throw ApiError{MPS_ERR_INVALID, "too many particles"};
        require_single("mps_upload_boundary");
        mesh.ok = false;
        nbobj = 1;
        alloc_boundary(uint32_t(n));
        std::vector<float4> hp(n), hv(n);
        for (size_t i = 0; i < n; i++) {
            hp[i] = make_float4(p[3 * i], p[3 * i + 1], p[3 * i + 2], r[i]);
            hv[i] = make_float4(0.f, 0.f, 0.f, m ? m[i] : 0.f);
        }
        B = uint32_t(n);
        if (n) {
            HIPCHK(hipMemcpyAsync(bpos.p, hp.data(), n * sizeof(float4), hipMemcpyHostToDevice, stream));
            HIPCHK(hipMemcpyAsync(bvel.p, hv.data(), n * sizeof(float4), hipMemcpyHostToDevice, stream));
            mps::launch::zero_words(reinterpret_cast<uint32_t*>(bcolor.p), 4 * n, stream);
            mps::launch::zero_words(reinterpret_cast<uint32_t*>(bvol.p), n, stream);
            mps::launch::zero_words(bface.p, n, stream);
            mps::launch::zero_words(reinterpret_cast<uint32_t*>(bbcc.p), 2 * n, stream);
            mps::launch::zero_words(cbb.p, n, stream);
        }
        bh.cells_init = false;
        bh.n_valid = 0;
        bnd_dirty = true;
        params_dirty = true;
        sync();
    }
    void sample_mesh(const double* verts, size_t nv, const uint32_t* tris, size_t nt, const double* center,
                     const double* size, const mps_mesh_material* m) {
        if (!verts || !tris || nv == 0 || nt == 0) throw ApiError{MPS_ERR_INVALID, "empty mesh"};
        require_single("mps_sample_boundary_mesh");
        params_dirty = true;
        nbobj = 1;
        for (size_t i = 0; i < 3 * nt; i++)
            if (tris[i] >= nv) throw ApiError{MPS_ERR_INVALID, "triangle index out of range"};
        std::vector<double> V(verts, verts + 3 * nv);
        if (center && size) {  // Translate with AABB::AddPoint's else-if (MPSAABB.h:20-28, MPSMeshObject.cu:130-151)
            double mn[3] = {1.7976931348623157e308, 1.7976931348623157e308, 1.7976931348623157e308};
            double mx[3] = {-1.7976931348623157e308, -1.7976931348623157e308, -1.7976931348623157e308};
            for (size_t i = 0; i < nv; i++)
                for (int k = 0; k < 3; k++) {
                    const double q = verts[3 * i + k];
                    if (mn[k] > q) mn[k] = q;
                    else if (mx[k] < q) mx[k] = q;
                }
            for (int k = 0; k < 3; k++) {
                const double oc = (mn[k] + mx[k]) * 0.5, os = mx[k] - mn[k];
                const double sc = os > 1.0e-10 ? size[k] / os : 0.0;
                for (size_t i = 0; i < nv; i++) V[3 * i + k] = (V[3 * i + k] - oc) * sc + center[k];
            }
        }
        double area = 0.0;  // InitFaceInfo (MPSMeshObject.cu:153-175)
        for (size_t t = 0; t < nt; t++) {
            const uint32_t* f = tris + 3 * t;
            const double ax = V[3 * f[1]] - V[3 * f[0]], ay = V[3 * f[1] + 1] - V[3 * f[0] + 1], az = V[3 * f[1] + 2] - V[3 * f[0] + 2];
            const double bx = V[3 * f[2]] - V[3 * f[0]], by = V[3 * f[2] + 1] - V[3 * f[0] + 1], bz = V[3 * f[2] + 2] - V[3 * f[0] + 2];
            const double cx = ay * bz - by * az, cy = az * bx - bz * ax, cz = ax * by - bx * ay;
            area += 0.5 * std::sqrt((cx * cx + cy * cy) + cz * cz);
        }
        std::vector<double> vmass(nv, area * m->density / double(nv));
        DBuf<double> dV, dmass, dpos;
        DBuf<uint32_t> dF, dR, dcnt, dface;
        dV.reserve(3 * nv); dmass.reserve(nv); dF.reserve(3 * nt); dR.reserve(nt); dcnt.reserve(nt + 1);
        HIPCHK(hipMemcpyAsync(dV.p, V.data(), 3 * nv * sizeof(double), hipMemcpyHostToDevice, stream));
        HIPCHK(hipMemcpyAsync(dmass.p, vmass.data(), nv * sizeof(double), hipMemcpyHostToDevice, stream));
        HIPCHK(hipMemcpyAsync(dF.p, tris, 3 * nt * sizeof(uint32_t), hipMemcpyHostToDevice, stream));
        {  // which face samples each shared vertex and edge (buildAdjacency + RTriangleBuild)
            const size_t slots = mps::launch::owner_table_slots(uint32_t(nt));
            DBuf<uint32_t> vmin, emin;
            DBuf<unsigned long long> ekey;
            vmin.reserve(nv); emin.reserve(slots); ekey.reserve(slots);
            mps::launch::owner_build(dF.p, uint32_t(nt), uint32_t(nv), vmin.p, ekey.p, emin.p, slots, dR.p, stream);
            sync();
            vmin.release(); emin.release(); ekey.release();
        }
        std::vector<uint32_t> rtri = down(dR.p, nt);
        DBuf<uint32_t> dshort;
        dshort.reserve(nt);
        mps::launch::sample_count(dV.p, dF.p, dR.p, uint32_t(nt), m->radius, dcnt.p, dshort.p, stream);
        std::vector<uint32_t> cnt(nt + 1, 0u), shorts(nt, 0u);
        HIPCHK(hipMemcpyAsync(cnt.data(), dcnt.p, nt * sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipMemcpyAsync(shorts.data(), dshort.p, nt * sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
        sync();
        dshort.release();
        std::vector<uint32_t> offs(nt + 1, 0u);
        for (size_t t = 0; t < nt; t++) offs[t + 1] = offs[t] + cnt[t];
        const uint32_t n = offs[nt];
        HIPCHK(hipMemcpyAsync(dcnt.p, offs.data(), (nt + 1) * sizeof(uint32_t), hipMemcpyHostToDevice, stream));
        dpos.reserve(3 * std::max<uint32_t>(n, 1));
        dface.reserve(std::max<uint32_t>(n, 1));
        mps::launch::zero_words(d_flags + 1, 1, stream);
        mps::launch::fill_u32(dface.p, std::max<uint32_t>(n, 1), 0xFFFFFFFFu, stream);
        mps::launch::sample_generate(dV.p, dF.p, dR.p, uint32_t(nt), m->radius, nullptr, dcnt.p, dpos.p, dface.p, d_flags + 1,
                                     stream);
        alloc_boundary(n);
        B = n;
        mps::launch::sample_finish(dV.p, dF.p, uint32_t(nt), dmass.p, n, dpos.p, dface.p, m->radius, 0, bpos.p, bvel.p, bbcc.p,
                                   bface.p, bcolor.p, stream);
        if (n) {
            mps::launch::zero_words(reinterpret_cast<uint32_t*>(bvol.p), n, stream);
            mps::launch::zero_words(cbb.p, n, stream);
        }
        HIPCHK(hipMemcpyAsync(h_flags + 1, d_flags + 1, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
        sync();
        dV.release(); dmass.release(); dpos.release(); dF.release(); dR.release(); dcnt.release(); dface.release();
        bh.cells_init = false;
        bh.n_valid = 0;
        bnd_dirty = true;
        if (h_flags[1]) throw ApiError{MPS_ERR_STATE, "sample count mismatch between count and generate passes"};
        mesh.V = std::move(V);
        mesh.F.assign(tris, tris + 3 * nt);
        mesh.R = std::move(rtri);
        mesh.vmass = std::move(vmass);
        mesh.count.assign(cnt.begin(), cnt.begin() + nt);
        mesh.shorts = std::move(shorts);
        mesh.ok = true;
    }

    // ParticleSampling on the same mesh after its vertices moved (IsSamplingParticleRequired +
    // GenerateBoundaryParticle_kernel + SetBarycentric_kernel, MPSParticleSamplingUtil.cu:14-133,
    // .cuh:69-294). The count pass runs on the GPU; the host compares each face's count and
    // shortest edge with the kept ones and scans the offsets (faces are few); faces that
    // changed are generated anew and finished on the GPU, the other faces' particles move to
    // their new offset in a per-slot copy. The arrays are resized by index first, as
    // BoundaryParticleObject::Resize does: volume always, and mass/velocity with faithful
    // quirks, stay with the index; otherwise mass and velocity move with the particle.
    void resample_mesh(const double* verts, size_t nv, const mps_mesh_material* m, size_t* regenerated) {
        require_single("mps_resample_boundary_mesh");
        if (!mesh.ok || nbobj != 1) throw ApiError{MPS_ERR_STATE, "no mesh sampled on this context (mps_sample_boundary_mesh first)"};
        if (!verts || 3 * nv != mesh.V.size()) throw ApiError{MPS_ERR_INVALID, "vertex count differs from the sampled mesh"};
        const size_t nt = mesh.count.size();
        std::vector<double> V(verts, verts + 3 * nv);
        DBuf<double> dV, dpos, dmass;
        DBuf<uint32_t> dF, dR, dcnt, dshort, dface, dcurr, dprev, p_face;
        DBuf<uint8_t> dgen;
        DBuf<float4> p_pos, p_col, p_vel;
        DBuf<float2> p_bcc;
        struct Release {  // scratch is freed on every way out (DBuf frees only on release())
            std::function<void()> f;
            ~Release() { f(); }
        } scratch{[&] {
            dV.release(); dpos.release(); dmass.release(); dF.release(); dR.release(); dcnt.release(); dshort.release();
            dface.release(); dcurr.release(); dprev.release(); p_face.release(); dgen.release(); p_pos.release();
            p_col.release(); p_vel.release(); p_bcc.release();
        }};
        dV.reserve(3 * nv); dF.reserve(3 * nt); dR.reserve(nt); dcnt.reserve(nt); dshort.reserve(nt);
        HIPCHK(hipMemcpyAsync(dV.p, V.data(), 3 * nv * sizeof(double), hipMemcpyHostToDevice, stream));
        HIPCHK(hipMemcpyAsync(dF.p, mesh.F.data(), 3 * nt * sizeof(uint32_t), hipMemcpyHostToDevice, stream));
        HIPCHK(hipMemcpyAsync(dR.p, mesh.R.data(), nt * sizeof(uint32_t), hipMemcpyHostToDevice, stream));
        mps::launch::sample_count(dV.p, dF.p, dR.p, uint32_t(nt), m->radius, dcnt.p, dshort.p, stream);
        std::vector<uint32_t> cnt(nt), shorts(nt);
        HIPCHK(hipMemcpyAsync(cnt.data(), dcnt.p, nt * sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipMemcpyAsync(shorts.data(), dshort.p, nt * sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
        sync();
        std::vector<uint32_t> prev(nt + 1, 0u), curr(nt + 1, 0u);
        std::vector<uint8_t> gen(nt, 0u);
        size_t ngen = 0;
        for (size_t t = 0; t < nt; t++) {
            prev[t + 1] = prev[t] + mesh.count[t];
            curr[t + 1] = curr[t] + cnt[t];
            if (shorts[t] != mesh.shorts[t] || cnt[t] != mesh.count[t]) {
                gen[t] = 1u;
                ngen++;
            }
        }
        mesh.V = std::move(V);
        mesh.count = cnt;
        mesh.shorts = shorts;
        if (regenerated) *regenerated = ngen;
        if (ngen == 0) return;
        drop_rollout();
        if (prev[nt] != B) throw ApiError{MPS_ERR_STATE, "boundary particles no longer match the sampled mesh"};
        const uint32_t B0 = B, n = curr[nt];
        params_dirty = true;
        // the previous particle arrays
        const size_t m0 = std::max<uint32_t>(B0, 1);
        p_pos.reserve(m0); p_col.reserve(m0); p_vel.reserve(m0); p_bcc.reserve(m0); p_face.reserve(m0);
        if (B0) {
            HIPCHK(hipMemcpyAsync(p_pos.p, bpos.p, B0 * sizeof(float4), hipMemcpyDeviceToDevice, stream));
            HIPCHK(hipMemcpyAsync(p_col.p, bcolor.p, B0 * sizeof(float4), hipMemcpyDeviceToDevice, stream));
            HIPCHK(hipMemcpyAsync(p_vel.p, bvel.p, B0 * sizeof(float4), hipMemcpyDeviceToDevice, stream));
            HIPCHK(hipMemcpyAsync(p_bcc.p, bbcc.p, B0 * sizeof(float2), hipMemcpyDeviceToDevice, stream));
            HIPCHK(hipMemcpyAsync(p_face.p, bface.p, B0 * sizeof(uint32_t), hipMemcpyDeviceToDevice, stream));
        }
        // resize by index: the leading min(B0, n) values stay, new slots are zero
        auto resize_keep = [&](auto& buf, size_t words_per) {
            using E = std::remove_reference_t<decltype(*buf.p)>;
            std::remove_reference_t<decltype(buf)> nb;
            nb.reserve(std::max<uint32_t>(n, 1));
            const size_t keep = std::min(B0, n);
            if (keep) HIPCHK(hipMemcpyAsync(nb.p, buf.p, keep * sizeof(E), hipMemcpyDeviceToDevice, stream));
            if (n > keep) mps::launch::zero_words(reinterpret_cast<uint32_t*>(nb.p + keep), (n - keep) * words_per, stream);
            sync();
            buf.release();
            buf = nb;
            nb.p = nullptr;
            nb.cap = 0;
        };
        resize_keep(bvel, 4);
        resize_keep(bvol, 1);
        bpos.reserve(std::max<uint32_t>(n, 1)); bcolor.reserve(std::max<uint32_t>(n, 1));
        bbcc.reserve(std::max<uint32_t>(n, 1)); bface.reserve(std::max<uint32_t>(n, 1));
        bS.reserve(std::max<uint32_t>(n, 1)); cbb.reserve(std::max<uint32_t>(n, 1));
        if (n) mps::launch::zero_words(cbb.p, n, stream);
        // faces that changed: generate and finish; the rest: copy
        dpos.reserve(3 * size_t(std::max<uint32_t>(n, 1))); dface.reserve(std::max<uint32_t>(n, 1));
        dmass.reserve(nv); dcurr.reserve(nt + 1); dprev.reserve(nt + 1); dgen.reserve(nt);
        HIPCHK(hipMemcpyAsync(dmass.p, mesh.vmass.data(), nv * sizeof(double), hipMemcpyHostToDevice, stream));
        HIPCHK(hipMemcpyAsync(dcurr.p, curr.data(), (nt + 1) * sizeof(uint32_t), hipMemcpyHostToDevice, stream));
        HIPCHK(hipMemcpyAsync(dprev.p, prev.data(), (nt + 1) * sizeof(uint32_t), hipMemcpyHostToDevice, stream));
        HIPCHK(hipMemcpyAsync(dgen.p, gen.data(), nt, hipMemcpyHostToDevice, stream));
        mps::launch::zero_words(d_flags + 1, 1, stream);
        mps::launch::fill_u32(dface.p, std::max<uint32_t>(n, 1), 0xFFFFFFFFu, stream);
        mps::launch::sample_generate(dV.p, dF.p, dR.p, uint32_t(nt), m->radius, dgen.p, dcurr.p, dpos.p, dface.p, d_flags + 1,
                                     stream);
        mps::launch::sample_finish(dV.p, dF.p, uint32_t(nt), dmass.p, n, dpos.p, dface.p, m->radius, 1, bpos.p, bvel.p, bbcc.p,
                                   bface.p, bcolor.p, stream);
        mps::launch::sample_copy(uint32_t(nt), dcurr.p, dprev.p, dgen.p, n, p_pos.p, p_bcc.p, p_face.p, p_col.p,
                                 cfg.faithful_quirks ? nullptr : p_vel.p, bpos.p, bbcc.p, bface.p, bcolor.p, bvel.p, stream);
        HIPCHK(hipMemcpyAsync(h_flags + 1, d_flags + 1, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
        sync();
        B = n;
        bh.cells_init = false;
        bh.n_valid = 0;
        bnd_dirty = true;
        if (h_flags[1]) throw ApiError{MPS_ERR_STATE, "sample count mismatch between count and generate passes"};
    }

    // ------------------------------------------------------------------ array access
    template <class T>
    std::vector<T> down(const T* p, size_t n) {
        std::vector<T> h(n);
        if (n) {
            HIPCHK(hipMemcpyAsync(h.data(), p, n * sizeof(T), hipMemcpyDeviceToHost, stream));
            sync();
        }
        return h;
    }
    template <class T>
    void up(T* p, const std::vector<T>& h) {
        if (!h.empty()) {
            HIPCHK(hipMemcpyAsync(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, stream));
            sync();
        }
    }
    // CSR in reference order from slot-order ELL lists; fluid neighbour slots are mapped to
    // reference indices (map_ids), boundary neighbours are already reference indices.
    std::vector<uint32_t> csr_idx_fluid(const uint32_t* cnt) {
        std::vector<uint32_t> c = down(cnt, N), iv = down(inv.p, Nref), idx(size_t(Nref) + 1, 0u);
        for (uint32_t i = 0; i < Nref; i++) idx[i + 1] = idx[i] + c[iv[i]];
        return idx;
    }
    std::vector<uint32_t> csr_nei_fluid(const uint32_t* nbrs, const uint32_t* cnt, uint32_t K, bool map_ids) {
        std::vector<uint32_t> c = down(cnt, N), iv = down(inv.p, Nref), pm = down(perm.p, N);
        std::vector<uint32_t> ell = down(nbrs, size_t(K) * N);
        std::vector<uint32_t> out;
        for (uint32_t i = 0; i < Nref; i++) {
            const uint32_t sl = iv[i];
            for (uint32_t k = 0; k < c[sl]; k++) {
                const uint32_t j = ell[size_t(k) * N + sl];
                out.push_back(map_ids ? pm[j] : j);
            }
        }
        return out;
    }
    std::vector<uint32_t> csr_nei_rows(const uint32_t* nbrs, const uint32_t* cnt, uint32_t n, uint32_t K) {
        std::vector<uint32_t> c = down(cnt, n);
        std::vector<uint32_t> rows = down(nbrs, size_t(K) * n);
        std::vector<uint32_t> out;
        for (uint32_t i = 0; i < n; i++)
            for (uint32_t k = 0; k < c[i]; k++) out.push_back(rows[size_t(i) * K + k]);
        return out;
    }
    std::vector<uint32_t> csr_idx(const uint32_t* cnt, uint32_t n) {
        std::vector<uint32_t> c = down(cnt, n), idx(size_t(n) + 1, 0u);
        for (uint32_t i = 0; i < n; i++) idx[i + 1] = idx[i] + c[i];
        return idx;
    }
    std::vector<uint32_t> csr_nei(const uint32_t* nbrs, const uint32_t* cnt, uint32_t n, uint32_t K) {
        std::vector<uint32_t> c = down(cnt, n);
        std::vector<uint32_t> ell = down(nbrs, size_t(K) * n);
        std::vector<uint32_t> out;
        for (uint32_t i = 0; i < n; i++)
            for (uint32_t k = 0; k < c[i]; k++) out.push_back(ell[size_t(k) * n + i]);
        return out;
    }
    // returns (count, elem_bytes) and optionally the data as bytes
    void array_io(int32_t id, size_t* count, size_t* eb, void* dst, const void* src, size_t bytes) {
        // fluid arrays live in slot order; the API shows reference order (out[perm[s]] = a[s])
        std::vector<uint32_t> hperm;
        auto get_perm = [&]() -> const std::vector<uint32_t>& {
            if (hperm.empty() && N) hperm = down(perm.p, N);
            return hperm;
        };
        auto fl4 = [&](DBuf<float4>& a, int comp /* -1 = xyz, 3 = w, 4 = all */) {
            const size_t w = comp == -1 ? 3 : (comp == 4 ? 4 : 1);
            *count = Nref;
            *eb = 4 * w;
            if (!(dst || src)) return;
            if (bytes != size_t(Nref) * 4 * w) throw ApiError{MPS_ERR_INVALID, "size mismatch"};
            const auto& P = get_perm();
            std::vector<float4> h = down(a.p, N);
            if (dst) {
                float* d = static_cast<float*>(dst);
                for (uint32_t s = 0; s < N; s++) {
                    if (P[s] == mps::dev::kEmpty) continue;
                    const float q[4] = {h[s].x, h[s].y, h[s].z, h[s].w};
                    if (comp == -1) for (int k = 0; k < 3; k++) d[3 * size_t(P[s]) + k] = q[k];
                    else if (comp == 4) for (int k = 0; k < 4; k++) d[4 * size_t(P[s]) + k] = q[k];
                    else d[P[s]] = q[3];
                }
            } else {
                const float* q = static_cast<const float*>(src);
                for (uint32_t s = 0; s < N; s++) {
                    if (P[s] == mps::dev::kEmpty) continue;
                    if (comp == -1) { h[s].x = q[3 * size_t(P[s])]; h[s].y = q[3 * size_t(P[s]) + 1]; h[s].z = q[3 * size_t(P[s]) + 2]; }
                    else if (comp == 4) { h[s].x = q[4 * size_t(P[s])]; h[s].y = q[4 * size_t(P[s]) + 1]; h[s].z = q[4 * size_t(P[s]) + 2]; h[s].w = q[4 * size_t(P[s]) + 3]; }
                    else h[s].w = q[P[s]];
                }
                up(a.p, h);
            }
        };
        auto fl1 = [&](DBuf<float>& a) {
            *count = Nref;
            *eb = 4;
            if (!(dst || src)) return;
            if (bytes != size_t(Nref) * 4) throw ApiError{MPS_ERR_INVALID, "size mismatch"};
            const auto& P = get_perm();
            std::vector<float> h = down(a.p, N);
            if (dst) {
                float* d = static_cast<float*>(dst);
                for (uint32_t s = 0; s < N; s++)
                    if (P[s] != mps::dev::kEmpty) d[P[s]] = h[s];
            } else {
                const float* q = static_cast<const float*>(src);
                for (uint32_t s = 0; s < N; s++)
                    if (P[s] != mps::dev::kEmpty) h[s] = q[P[s]];
                up(a.p, h);
            }
        };
        auto f4xyz = [&](DBuf<float4>& a, uint32_t n) {
            *count = n;
            *eb = 12;
            if (dst || src) {
                if (bytes != size_t(n) * 12) throw ApiError{MPS_ERR_INVALID, "size mismatch"};
                std::vector<float4> h = down(a.p, n);
                if (dst) {
                    float* d = static_cast<float*>(dst);
                    for (uint32_t i = 0; i < n; i++) { d[3 * i] = h[i].x; d[3 * i + 1] = h[i].y; d[3 * i + 2] = h[i].z; }
                } else {
                    const float* s = static_cast<const float*>(src);
                    for (uint32_t i = 0; i < n; i++) { h[i].x = s[3 * i]; h[i].y = s[3 * i + 1]; h[i].z = s[3 * i + 2]; }
                    up(a.p, h);
                }
            }
        };
        auto f4w = [&](DBuf<float4>& a, uint32_t n) {
            *count = n;
            *eb = 4;
            if (dst || src) {
                if (bytes != size_t(n) * 4) throw ApiError{MPS_ERR_INVALID, "size mismatch"};
                std::vector<float4> h = down(a.p, n);
                if (dst) {
                    float* d = static_cast<float*>(dst);
                    for (uint32_t i = 0; i < n; i++) d[i] = h[i].w;
                } else {
                    const float* s = static_cast<const float*>(src);
                    for (uint32_t i = 0; i < n; i++) h[i].w = s[i];
                    up(a.p, h);
                }
            }
        };
        auto raw = [&](auto* p, size_t n, size_t e) {
            *count = n;
            *eb = e;
            if (dst || src) {
                if (bytes != n * e) throw ApiError{MPS_ERR_INVALID, "size mismatch"};
                if (n == 0) return;
                if (dst) HIPCHK(hipMemcpyAsync(dst, p, bytes, hipMemcpyDeviceToHost, stream));
                else HIPCHK(hipMemcpyAsync(p, src, bytes, hipMemcpyHostToDevice, stream));
                sync();
            }
        };
        auto ro = [&](const std::vector<uint32_t>& v) {
            *count = v.size();
            *eb = 4;
            if (src) throw ApiError{MPS_ERR_INVALID, "array is read-only"};
            if (dst) {
                if (bytes != v.size() * 4) throw ApiError{MPS_ERR_INVALID, "size mismatch"};
                std::memcpy(dst, v.data(), bytes);
            }
        };
        auto ro_raw = [&](const uint32_t* p, size_t n) {
            if (src) throw ApiError{MPS_ERR_INVALID, "array is read-only"};
            raw(const_cast<uint32_t*>(p), n, 4);
        };
        const bool hash_id = id >= MPS_ARR_FHASH_KEY && id <= MPS_ARR_BHASH_END;
        if (hash_id && S > 1) throw ApiError{MPS_ERR_STATE, "hash tables are exported for a single scene only"};
        switch (id) {
            case MPS_ARR_FLUID_POS:
                fl4(pos, -1);
                if (src) pairs_valid = halo_valid = false;
                break;
            case MPS_ARR_FLUID_VEL: fl4(vel, -1); break;
            case MPS_ARR_FLUID_RADIUS:
                fl4(pos, 3);
                if (src) pairs_valid = halo_valid = false;
                break;
            case MPS_ARR_FLUID_MASS: fl4(vel, 3); break;
            case MPS_ARR_FLUID_COLOR: fl4(color, 4); break;
            case MPS_ARR_FLUID_DENSITY: fl1(rho); break;
            case MPS_ARR_FLUID_ALPHA: fl1(alpha); break;
            case MPS_ARR_FLUID_KAPPA: fl1(kappa); break;
            case MPS_ARR_FLUID_DELTA: fl1(delta); break;
            case MPS_ARR_FLUID_FACTOR_ST: fl1(fst); break;
            case MPS_ARR_FLUID_CG_B: fl4(cgb, -1); break;
            case MPS_ARR_BND_POS:
                f4xyz(bpos, B);
                if (src) bnd_dirty = true;
                break;
            case MPS_ARR_BND_RADIUS:
                f4w(bpos, B);
                if (src) bnd_dirty = true;
                break;
            case MPS_ARR_BND_VOLUME: raw(bvol.p, B, 4); break;
            case MPS_ARR_BND_MASS: f4w(bvel, B); break;
            case MPS_ARR_BND_VEL: f4xyz(bvel, B); break;
            case MPS_ARR_BND_FACE_ID: raw(bface.p, B, 4); break;
            case MPS_ARR_BND_BCC: raw(bbcc.p, B, 8); break;
            case MPS_ARR_BND_OBJECT:
                if (nbobj > 1) ro_raw(bobj.p, B);
                else ro(std::vector<uint32_t>(B, 0u));
                break;
            case MPS_ARR_FHASH_KEY: {
                std::vector<uint32_t> v;
                if (dst) fluid_hash_export(&v, nullptr); else v.resize(fh.n_valid);
                ro(v);
                break;
            }
            case MPS_ARR_FHASH_ID: {
                std::vector<uint32_t> v;
                if (dst) fluid_hash_export(nullptr, &v); else v.resize(fh.n_valid);
                ro(v);
                break;
            }
            case MPS_ARR_FHASH_START: ro(dst ? fluid_cells_export(true) : std::vector<uint32_t>(fh.cells_init ? cells() : 0)); break;
            case MPS_ARR_FHASH_END: ro(dst ? fluid_cells_export(false) : std::vector<uint32_t>(fh.cells_init ? cells() : 0)); break;
            case MPS_ARR_BHASH_KEY: ro_raw(bh.key.p, bh.n_valid); break;
            case MPS_ARR_BHASH_ID: ro_raw(bh.id.p, bh.n_valid); break;
            case MPS_ARR_BHASH_START: ro_raw(bh.start.p, bh.cells_init ? cells() : 0); break;
            case MPS_ARR_BHASH_END: ro_raw(bh.end.p, bh.cells_init ? cells() : 0); break;
            case MPS_ARR_NEI_FF_IDX: ro(csr_idx_fluid(cff.p)); break;
            case MPS_ARR_NEI_FF: ro(csr_nei_fluid(nff.p, cff.p, Kff, true)); break;
            case MPS_ARR_NEI_FB_IDX: ro(csr_idx_fluid(cfb.p)); break;
            case MPS_ARR_NEI_FB: ro(csr_nei_fluid(nfb.p, cfb.p, Kfb, false)); break;
            case MPS_ARR_NEI_BB_IDX: ro(csr_idx(cbb.p, B)); break;
            case MPS_ARR_NEI_BB: ro(csr_nei_rows(nbb.p, cbb.p, B, Kbb)); break;
            default: throw ApiError{MPS_ERR_INVALID, "unknown array id"};
        }
    }

    ~mps_ctx() {
        pos.release(); vel.release(); color.release(); cgb.release(); cgx.release(); cgr.release();
        v3x.release(); v3r.release(); v3Ap.release(); v3p.release(); v3q.release();
        cgAp.release(); jprev.release(); for (auto& t : jt) t.release(); jcand.release(); tmp4.release();
        rho.release(); alpha.release(); kappa.release(); delta.release(); fst.release(); tmp1.release();
        pos_alt.release(); vel_alt.release(); color_alt.release(); cgb_alt.release(); rho_alt.release();
        alpha_alt.release(); kappa_alt.release(); delta_alt.release(); fst_alt.release();
        bpos.release(); bvel.release(); bcolor.release(); bvol.release(); bS.release(); bface.release(); bbcc.release();
        fh.release(); bh.release();
        zkey.release(); zid.release(); zkey2.release(); zid2.release(); ztemp.release(); tmpu.release(); tmp2.release();
        nff.release(); cff.release(); nfb.release(); cfb.release(); nbb.release(); cbb.release(); nfr.release(); cfr.release();
        bobj.release(); bS4.release();
        partials.release();
        arrive.release();
        pairG_ff.release(); pairG_fb.release(); pairAB.release(); pairA.release(); irho.release(); halo_idx.release(); halo_n.release(); nb16.release(); nstat.release(); cgrecA.release(); cgrecB.release();
        if (copy_stream) (void)hipStreamSynchronize(copy_stream);
        for (ReadbackSlot& sl : rb) {
            sl.buf.release();
            if (sl.packed) (void)hipEventDestroy(sl.packed);
            if (sl.copied) (void)hipEventDestroy(sl.copied);
        }
        if (copy_stream) (void)hipStreamDestroy(copy_stream);
        perm.release(); inv.release(); keyref.release(); vals.release(); src.release(); lkey.release(); hash_perm.release();
        d_params.release(); blk_scene.release(); bscene.release(); d_delta.release(); mkey.release();
        clear_graphs();
        for (auto& t : tape) t.release();
        for (auto& c : ckpts) c.release();
        ck_end.release();
        perm0.release(); d_L.release(); rev.release(); gobs.release(); adx.release(); adv.release(); adx2.release(); adv2.release();
        aux4.release(); gvec.release(); avin.release(); abar.release(); rbar.release(); dbar.release(); fbar.release();
        fstr.release(); cabT.release(); acc_mu.release(); acc_sig.release(); lam.release(); lam2.release();
        if (d_active) (void)hipFree(d_active);
        if (h_active) (void)hipHostFree(h_active);
        for (const Rec& r : recs) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
        for (hipEvent_t e : evpool) (void)hipEventDestroy(e);
        if (d_ctr) (void)hipFree(d_ctr);
        if (d_solver) (void)hipFree(d_solver);
        if (d_flags) (void)hipFree(d_flags);
        if (d_grid) (void)hipFree(d_grid);
        if (d_err3) (void)hipFree(d_err3);
        if (h_ctr) (void)hipHostFree(h_ctr);
        if (h_solver) (void)hipHostFree(h_solver);
        if (h_flags) (void)hipHostFree(h_flags);
        if (h_err3) (void)hipHostFree(h_err3);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {
template <class F>
int guarded(mps_ctx* c, F&& f) {
    if (!c) return MPS_ERR_INVALID;
    try {
        HIPCHK(hipSetDevice(c->device));
        f();
        c->err.clear();
        return MPS_OK;
    } catch (const HipError& e) {
        c->err = std::string(e.what) + ": " + hipGetErrorString(e.e);
        return MPS_ERR_HIP;
    } catch (const mps::launch::LaunchError& e) {
        c->err = std::string("kernel launch: ") + hipGetErrorString(e.e);
        return MPS_ERR_HIP;
    } catch (const ApiError& e) {
        c->err = e.msg;
        return e.code;
    } catch (const std::bad_alloc&) {
        c->err = "host allocation failed";
        return MPS_ERR_NOMEM;
    } catch (const std::exception& e) {  // nothing may cross the C ABI
        c->err = std::string("internal error: ") + e.what();
        return MPS_ERR_STATE;
    } catch (...) {
        c->err = "internal error";
        return MPS_ERR_STATE;
    }
}

// Wavefront OBJ reading as MeshObject::LoadMesh does it (MPSMeshObject.cu:58-103): the file is
// a stream of whitespace-separated tokens; a token longer than one character is skipped (so
// `vn`, `vt`, `usemtl`, ... and every word of a comment are ignored, but a one-character word
// is not); `v` is followed by three reals, `f` by three 1-based vertex indices (a fourth index
// of a quad is a one-character token or longer and is skipped like any other, so a quad keeps
// its first triangle). Where the reference's stream would stop making progress or read garbage
// (a malformed number, an index of 0 or past the vertices, a missing file, no faces or no
// vertices), this reader reports an error instead. Indices written `a/b/c` use `a`.
struct ObjMesh {
    std::vector<double> v;
    std::vector<uint32_t> f;
};
ObjMesh read_obj_file(const char* path) {
    if (!path) throw ApiError{MPS_ERR_INVALID, "OBJ path is null"};
    FILE* fp = std::fopen(path, "rb");
    if (!fp) throw ApiError{MPS_ERR_INVALID, std::string("cannot open OBJ file ") + path};
    std::string text;
    char buf[1 << 16];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof(buf), fp)) > 0) text.append(buf, got);
    std::fclose(fp);
    size_t at = 0;
    auto next = [&](std::string& tok) {
        while (at < text.size() && std::isspace(static_cast<unsigned char>(text[at]))) at++;
        const size_t b = at;
        while (at < text.size() && !std::isspace(static_cast<unsigned char>(text[at]))) at++;
        tok.assign(text, b, at - b);
        return !tok.empty();
    };
    ObjMesh m;
    std::string tok;
    std::vector<uint64_t> idx;
    while (next(tok)) {
        if (tok.size() > 1) continue;
        if (tok[0] == 'v') {
            for (int k = 0; k < 3; k++) {
                std::string t;
                char* end = nullptr;
                const double x = next(t) ? std::strtod(t.c_str(), &end) : 0.0;
                if (t.empty() || !end || *end != '\0') throw ApiError{MPS_ERR_INVALID, "malformed vertex in OBJ file"};
                m.v.push_back(x);
            }
        } else if (tok[0] == 'f') {
            for (int k = 0; k < 3; k++) {
                std::string t;
                if (!next(t)) throw ApiError{MPS_ERR_INVALID, "malformed face in OBJ file"};
                const size_t slash = t.find('/');
                const std::string a = t.substr(0, slash);
                if (a.empty() || a.size() > 10 || a.find_first_not_of("0123456789") != std::string::npos)
                    throw ApiError{MPS_ERR_INVALID, "malformed face index in OBJ file: " + t};
                idx.push_back(std::stoull(a));
            }
        }
    }
    const size_t nv = m.v.size() / 3;
    if (nv == 0 || idx.empty()) throw ApiError{MPS_ERR_INVALID, "OBJ file has no vertices or no faces"};
    m.f.reserve(idx.size());
    for (const uint64_t q : idx) {
        if (q == 0 || q > nv) throw ApiError{MPS_ERR_INVALID, "OBJ face index out of range"};
        m.f.push_back(static_cast<uint32_t>(q - 1));
    }
    return m;
}
}  // namespace

extern "C" {

const char* mps_version(void) { return "mps_sph 0.1 (gfx950)"; }

int mps_config_init(mps_config* cfg) {
    if (!cfg) return MPS_ERR_INVALID;
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->kernel = MPS_KERNEL_POLY6_SPIKY;       // USE_CUBIC_KERNEL 0 (MPSSPHKernel.cuh:5)
    cfg->visc_solver = MPS_VISC_CG;             // MPSSystem.cpp:266
    cfg->faithful_quirks = 1;
    cfg->device = 0;
    cfg->hash_size[0] = cfg->hash_size[1] = cfg->hash_size[2] = 64;  // MPSSystem.cpp:72
    cfg->cell_size = 0.0;                       // = material radius (MPSSystem.cpp:71)
    cfg->df_tol = 1.0e-3;                       // MPSSPHDFSPHUtil.cu:419
    cfg->df_min_iter = 1;
    cfg->cd_tol = 1.0e-4;                       // MPSSPHDFSPHUtil.cu:333
    cfg->cd_min_iter = 2;
    cfg->max_iter = 100;                        // MPSSPHDFSPHUtil.cu:299,405
    cfg->cg_tol = 1.0e-2;                       // MPSSPHNonPressureUtil.cu:159
    cfg->cg_max_iter = 100;                     // MPSSolver.cuh:31
    cfg->cg_factor = 0.01;                      // MPSSPHNonPressureUtil.cu:166-169
    cfg->jacobi_max_iter = 100;                 // SPHUtil.cu:373
    cfg->jacobi_tol = 1.0e-4;                   // SPHUtil.cu:436-437
    cfg->zsort_interval = 50;                   // MPSSystem.cpp:220
    cfg->nei_capacity = 64;
    cfg->host_sync_every = 0;
    cfg->adj_cg_max_iter = 300;
    cfg->adj_cg_tol = -1.0e-12;
    cfg->cg_fold_in_kernel = 0;
    cfg->cg_halo_depth = 0;
    cfg->cg_stream_coeffs = 1;  // measured faster on the 1M dam break (recomputing b_ij is VALU-bound)
    return MPS_OK;
}

int mps_create(const mps_config* cfg, mps_ctx** out) {
    if (!cfg || !out) return MPS_ERR_INVALID;
    for (int k = 0; k < 3; k++)
        if (cfg->hash_size[k] == 0 || (cfg->hash_size[k] & (cfg->hash_size[k] - 1))) return MPS_ERR_INVALID;
    if (size_t(cfg->hash_size[0]) * cfg->hash_size[1] * cfg->hash_size[2] > (size_t(1) << 31)) return MPS_ERR_INVALID;
    if (cfg->cg_halo_depth > 2) return MPS_ERR_INVALID;
    *out = nullptr;
    mps_ctx* c = new (std::nothrow) mps_ctx;
    if (!c) return MPS_ERR_NOMEM;
    c->cfg = *cfg;
    c->device = cfg->device;
    c->mat = mps_sph_material{0.01, 1.0, 0.025, 0.1, 100.0};  // SPHMaterial() defaults (MPSSPHMaterial.cpp:4-11)
    c->mV = c->mat.radius * c->mat.radius * c->mat.radius / 48.0 * 3.141592;
    c->mMass = c->mat.density * c->mV;
    const int rc = guarded(c, [&] {
        HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        HIPCHK(hipMalloc(reinterpret_cast<void**>(&c->d_ctr), sizeof(NbrCounters)));
        HIPCHK(hipMalloc(reinterpret_cast<void**>(&c->d_active), 4 * sizeof(uint32_t)));
        HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&c->h_active), 4 * sizeof(uint32_t)));
        HIPCHK(hipMalloc(reinterpret_cast<void**>(&c->d_flags), 8 * sizeof(uint32_t)));
        HIPCHK(hipMalloc(reinterpret_cast<void**>(&c->d_grid), sizeof(mps::GridSync)));
        HIPCHK(hipMalloc(reinterpret_cast<void**>(&c->d_err3), 4 * sizeof(double)));
        HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&c->h_ctr), sizeof(NbrCounters)));
        HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&c->h_flags), 8 * sizeof(uint32_t)));
        HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&c->h_err3), 4 * sizeof(double)));
        std::memset(c->h_flags, 0, 8 * sizeof(uint32_t));
        mps::launch::zero_words(c->d_flags, 8, c->stream);
        c->ensure_solvers();
        (void)scratch_address(c->device);  // once per device, with the device current
        c->sync();
    });
    if (rc != MPS_OK) {
        delete c;
        return rc;
    }
    *out = c;
    return MPS_OK;
}

void mps_destroy(mps_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    delete c;
}

const char* mps_last_error(const mps_ctx* c) { return c ? c->err.c_str() : "null context"; }

int mps_set_physics(mps_ctx* c, const mps_physics* p) {
    if (!c || !p) return MPS_ERR_INVALID;
    if (!(p->dt > 0.0)) {
        c->err = "dt must be positive";
        return MPS_ERR_INVALID;
    }
    c->phys = *p;
    for (auto& h : c->scenes) h.phys = *p;
    c->params_dirty = true;
    c->drop_rollout();
    return MPS_OK;
}

// SPHMaterial::SetParam (MPSSPHMaterial.cpp:13-19): volume = R^3/48*3.141592, mass = rho0*volume
int mps_set_fluid_material(mps_ctx* c, const mps_sph_material* m) {
    if (!c || !m) return MPS_ERR_INVALID;
    if (!(m->radius > 0.0) || !(m->density > 0.0)) {
        c->err = "material radius and density must be positive";
        return MPS_ERR_INVALID;
    }
    c->mat = *m;
    c->mV = m->radius * m->radius * m->radius / 48.0 * 3.141592;
    c->mMass = m->density * c->mV;
    for (auto& h : c->scenes) {
        h.mat = *m;
        h.V = c->mV;
        h.mass = c->mMass;
    }
    c->params_dirty = true;
    c->drop_rollout();
    return MPS_OK;
}

int mps_upload_fluid(mps_ctx* c, size_t n, const float* pos, const float* vel, const float* radius, const float* mass) {
    if (!c || (n && (!pos || !radius))) return MPS_ERR_INVALID;
    return guarded(c, [&] { c->upload_fluid(n, pos, vel, radius, mass); });
}

int mps_upload_boundary(mps_ctx* c, size_t n, const float* pos, const float* radius, const float* mass) {
    if (!c || (n && (!pos || !radius))) return MPS_ERR_INVALID;
    return guarded(c, [&] { c->upload_boundary(n, pos, radius, mass); });
}

int mps_sample_boundary_mesh(mps_ctx* c, const double* verts, size_t nv, const uint32_t* tris, size_t nt,
                             const double* center, const double* size, const mps_mesh_material* m) {
    if (!c || !m) return MPS_ERR_INVALID;
    return guarded(c, [&] { c->sample_mesh(verts, nv, tris, nt, center, size, m); });
}

int mps_set_boundary_objects(mps_ctx* c, uint32_t k, const uint32_t* counts) {
    if (!c) return MPS_ERR_INVALID;
    return guarded(c, [&] { c->set_boundary_objects(k, counts); });
}

int mps_resample_boundary_mesh(mps_ctx* c, const double* verts, size_t nv, const mps_mesh_material* m, size_t* regenerated) {
    if (!c || !m) return MPS_ERR_INVALID;
    return guarded(c, [&] { c->resample_mesh(verts, nv, m, regenerated); });
}

int mps_read_obj(mps_ctx* c, const char* path, double* verts, size_t* nv, uint32_t* tris, size_t* nt) {
    if (!nv || !nt) return MPS_ERR_INVALID;
    auto body = [&] {
        const ObjMesh m = read_obj_file(path);
        const size_t v = m.v.size() / 3, t = m.f.size() / 3;
        if (verts || tris) {
            if (!verts || !tris || *nv < v || *nt < t) throw ApiError{MPS_ERR_INVALID, "OBJ output buffers too small"};
            std::memcpy(verts, m.v.data(), m.v.size() * sizeof(double));
            std::memcpy(tris, m.f.data(), m.f.size() * sizeof(uint32_t));
        }
        *nv = v;
        *nt = t;
    };
    if (c) return guarded(c, body);
    try {  // no context: host work only, status code without a message
        body();
        return MPS_OK;
    } catch (const ApiError& e) {
        return e.code;
    } catch (const std::exception&) {
        return MPS_ERR_INVALID;
    }
}

int mps_sample_boundary_obj(mps_ctx* c, const char* path, const double* center, const double* size, const mps_mesh_material* m) {
    if (!c || !m) return MPS_ERR_INVALID;
    return guarded(c, [&] {
        const ObjMesh o = read_obj_file(path);
        c->sample_mesh(o.v.data(), o.v.size() / 3, o.f.data(), o.f.size() / 3, center, size, m);
    });
}

int mps_step(mps_ctx* c, mps_step_stats* out) {
    if (!c) return MPS_ERR_INVALID;
    int rc = guarded(c, [&] { c->step(out); });
    if (rc == MPS_OK && out && out->nonfinite) {
        c->err = "non-finite position or velocity after the step";
        rc = MPS_ERR_NONFINITE;
    }
    return rc;
}

int mps_step_noalloc(mps_ctx* c, mps_step_stats* out) {
    if (!c) return MPS_ERR_INVALID;
    int rc = guarded(c, [&] { c->step_noalloc(out); });
    if (rc == MPS_OK && out && out->nonfinite) {
        c->err = "non-finite position or velocity after the step";
        rc = MPS_ERR_NONFINITE;
    }
    return rc;
}

int mps_run_phase(mps_ctx* c, int32_t phase, mps_step_stats* out) {
    if (!c) return MPS_ERR_INVALID;
    return guarded(c, [&] {
        c->run_phase(phase, out);
        c->sync();
        if (c->prof_on) c->prof_collect();
    });
}

int mps_synchronize(mps_ctx* c) {
    return guarded(c, [&] {
        c->sync();
        if (c->copy_stream) HIPCHK(hipStreamSynchronize(c->copy_stream));  // readbacks in flight too
    });
}

int mps_array_info(mps_ctx* c, int32_t id, size_t* count, size_t* elem_bytes) {
    size_t n = 0, e = 0;
    const int rc = guarded(c, [&] { c->array_io(id, &n, &e, nullptr, nullptr, 0); });
    if (count) *count = n;
    if (elem_bytes) *elem_bytes = e;
    return rc;
}

int mps_get_array(mps_ctx* c, int32_t id, void* dst, size_t bytes) {
    if (!c || (!dst && bytes)) return MPS_ERR_INVALID;
    size_t n, e;
    return guarded(c, [&] {
        if (bytes == 0) {
            c->array_io(id, &n, &e, nullptr, nullptr, 0);
            if (n * e != 0) throw ApiError{MPS_ERR_INVALID, "size mismatch"};
            return;
        }
        c->array_io(id, &n, &e, dst, nullptr, bytes);
    });
}

int mps_set_array(mps_ctx* c, int32_t id, const void* src, size_t bytes) {
    if (!c || (!src && bytes)) return MPS_ERR_INVALID;
    size_t n, e;
    return guarded(c, [&] {
        if (bytes == 0) {
            c->array_io(id, &n, &e, nullptr, nullptr, 0);
            if (n * e != 0) throw ApiError{MPS_ERR_INVALID, "size mismatch"};
            return;
        }
        c->array_io(id, &n, &e, nullptr, src, bytes);
    });
}

int mps_get_frame(mps_ctx* c, uint32_t* f) {
    if (!c || !f) return MPS_ERR_INVALID;
    *f = c->frame;
    return MPS_OK;
}

int mps_set_frame(mps_ctx* c, uint32_t f) {
    if (!c) return MPS_ERR_INVALID;
    c->frame = f;
    return MPS_OK;
}

int mps_readback_async(mps_ctx* c, float* host_pos, float* host_radius, float* host_rgba, mps_fence* fence) {
    if (!c || !fence) return MPS_ERR_INVALID;
    return guarded(c, [&] { c->readback(host_pos, host_radius, host_rgba, false, fence); });
}

int mps_readback_async_f64(mps_ctx* c, double* host_pos, double* host_radius, float* host_rgba, mps_fence* fence) {
    if (!c || !fence) return MPS_ERR_INVALID;
    return guarded(c, [&] { c->readback(host_pos, host_radius, host_rgba, true, fence); });
}

int mps_fence_wait(mps_ctx* c, mps_fence* fence) {
    if (!c || !fence) return MPS_ERR_INVALID;
    return guarded(c, [&] {
        if (!fence->event) return;
        // the slot's event: this copy's, or a later one on the same in-order copy stream
        HIPCHK(hipEventSynchronize(static_cast<hipEvent_t>(fence->event)));
        fence->event = nullptr;
    });
}

int mps_profile_enable(mps_ctx* c, int32_t on) {
    return guarded(c, [&] {
        c->sync();
        c->prof_reset();
        c->prof_on = on != 0;
    });
}

int mps_calibration_copy(mps_ctx* c, size_t bytes, int32_t reps) {
    if (!c || reps < 0) return MPS_ERR_INVALID;
    return guarded(c, [&] {
        DBuf<float4> a, b;
        const size_t n = bytes / sizeof(float4);
        a.reserve(std::max<size_t>(n, 1));
        b.reserve(std::max<size_t>(n, 1));
        mps::launch::zero_words(reinterpret_cast<uint32_t*>(a.p), 4 * n, c->stream);
        for (int32_t r = 0; r < reps; r++) c->T_(mps_ctx::K_CALIB, [&] { mps::launch::calib_copy(a.p, b.p, n, c->stream); });
        for (int32_t r = 0; r < reps; r++) c->T_(mps_ctx::K_STREAM, [&] { mps::launch::stream_copy(a.p, b.p, n, c->stream); });
        c->sync();
        if (c->prof_on) c->prof_collect();
        a.release();
        b.release();
    });
}

int mps_profile_read(mps_ctx* c, mps_kernel_stat* out, size_t capacity, size_t* count) {
    if (!c || !count || (capacity && !out)) return MPS_ERR_INVALID;
    return guarded(c, [&] {
        c->sync();
        c->prof_collect();
        size_t n = 0;
        for (int k = 0; k < mps_ctx::K_COUNT; k++) {
            if (c->kn[k] == 0) continue;
            if (n < capacity) {
                mps_kernel_stat& st = out[n];
                std::memset(&st, 0, sizeof(st));
                std::strncpy(st.name, mps_ctx::kname(k), sizeof(st.name) - 1);
                st.launches = c->kn[k];
                st.useful = c->kuse[k];
                st.total_ms = c->kms[k];
            }
            n++;
        }
        *count = n;
    });
}

int mps_batch_clear(mps_ctx* c) {
    return guarded(c, [&] { c->batch_clear(); });
}

int mps_batch_add_scene(mps_ctx* c, const mps_physics* phys, const mps_sph_material* mat, size_t n_fluid, const float* pos,
                        const float* vel, const float* radius, const float* mass, size_t n_boundary, const float* bpos,
                        const float* bradius, const float* bmass) {
    if (!c || !phys || !mat || (n_fluid && (!pos || !radius)) || (n_boundary && (!bpos || !bradius))) return MPS_ERR_INVALID;
    return guarded(c, [&] {
        c->batch_add(*phys, *mat, n_fluid, pos, vel, radius, mass, n_boundary, bpos, bradius, bmass);
    });
}

int mps_batch_commit(mps_ctx* c) {
    return guarded(c, [&] { c->batch_commit(); });
}

int mps_scene_count(mps_ctx* c, uint32_t* n) {
    if (!c || !n) return MPS_ERR_INVALID;
    *n = c->scenes.empty() ? 1u : uint32_t(c->scenes.size());
    return MPS_OK;
}

int mps_scene_info(mps_ctx* c, uint32_t scene, mps_scene_range* out) {
    if (!c || !out) return MPS_ERR_INVALID;
    const std::vector<mps_ctx::SceneHost> sl = c->scene_list();
    if (scene >= sl.size()) return MPS_ERR_INVALID;
    out->fluid_offset = sl[scene].ref0;
    out->n_fluid = sl[scene].nf;
    out->boundary_offset = sl[scene].b0;
    out->n_boundary = sl[scene].nb;
    return MPS_OK;
}

int mps_scene_stats(mps_ctx* c, uint32_t scene, mps_step_stats* out) {
    if (!c || !out) return MPS_ERR_INVALID;
    if (scene >= c->sstats.size()) {
        c->err = "no statistics for this scene (no step yet?)";
        return MPS_ERR_STATE;
    }
    *out = c->sstats[scene];
    return MPS_OK;
}

int mps_scene_set_material(mps_ctx* c, uint32_t scene, const mps_sph_material* m) {
    if (!c || !m) return MPS_ERR_INVALID;
    if (c->scenes.empty()) return scene == 0 ? mps_set_fluid_material(c, m) : MPS_ERR_INVALID;
    if (scene >= c->scenes.size() || !(m->radius > 0.0) || !(m->density > 0.0)) return MPS_ERR_INVALID;
    mps_ctx::SceneHost& h = c->scenes[scene];
    h.mat = *m;
    h.V = m->radius * m->radius * m->radius / 48.0 * 3.141592;
    h.mass = m->density * h.V;
    c->params_dirty = true;
    c->drop_rollout();
    return MPS_OK;
}

int mps_autotune(mps_ctx* c, uint32_t steps, mps_tune_result* out) {
    return guarded(c, [&] { c->autotune(steps, out); });
}

int mps_adj_rollout(mps_ctx* c, uint32_t n_steps, uint32_t checkpoint_every, uint32_t* loop_counts) {
    return guarded(c, [&] { c->adj_rollout(n_steps, checkpoint_every, loop_counts); });
}

int mps_adj_rollout_traj(mps_ctx* c, uint32_t n_steps, uint32_t checkpoint_every, uint32_t n_obs, const uint32_t* obs_step,
                         float* pos_out, float* vel_out, uint32_t* loop_counts) {
    return guarded(c, [&] { c->adj_rollout(n_steps, checkpoint_every, loop_counts, n_obs, obs_step, pos_out, vel_out); });
}

int mps_adj_backward(mps_ctx* c, const float* dloss_dpos, const float* dloss_dvel, float* dloss_dvel0, double* dloss_dviscosity,
                     double* dloss_dsurface_tension) {
    if (!c || !dloss_dpos || !dloss_dvel0) return MPS_ERR_INVALID;
    return guarded(c, [&] {
        const uint32_t last = c->adj_n;
        c->adj_backward(1, &last, dloss_dpos, dloss_dvel, dloss_dvel0, dloss_dviscosity, dloss_dsurface_tension);
    });
}

int mps_adj_backward_traj(mps_ctx* c, uint32_t n_obs, const uint32_t* obs_step, const float* dloss_dpos,
                          const float* dloss_dvel, float* dloss_dvel0, double* dloss_dviscosity,
                          double* dloss_dsurface_tension) {
    if (!c || !obs_step || !dloss_dpos || !dloss_dvel0) return MPS_ERR_INVALID;
    return guarded(c, [&] {
        c->adj_backward(n_obs, obs_step, dloss_dpos, dloss_dvel, dloss_dvel0, dloss_dviscosity, dloss_dsurface_tension);
    });
}

int mps_host_alloc(mps_ctx* c, size_t bytes, void** out) {
    if (!c || !out) return MPS_ERR_INVALID;
    return guarded(c, [&] { HIPCHK(hipHostMalloc(out, bytes)); });
}

int mps_host_free(mps_ctx* c, void* p) {
    return guarded(c, [&] {
        if (p) HIPCHK(hipHostFree(p));
    });
}

}  // extern "C"
