// hash.hip — spatial hash and neighbour search for gfx950.
//
// Replaces SpatialHash::UpdateHash / ZSort / BuildNeighorhood (MPSSpatialHash.cu:53-218,
// MPSSpatialHash.cuh:83-271). Design:
//   * cell keys computed in double exactly like GetGridPos (trunc(x*(1/cell)) + H/2);
//   * a stable LSD radix sort written here (8-bit digits, wave64 ballot ranking), so ties
//     keep ascending particle id — the order the reference's neighbour lists depend on;
//   * cell start/end tables cleared incrementally (only cells occupied last step), plus a
//     copy of the positions in key order so the 27-cell scan reads contiguous memory;
//   * one search pass per particle (no count/scan/fill double scan): fluid neighbour ids go
//     to column-major ELL storage nbr[k*N + i] (a wave reads 256 contiguous bytes per k),
//     the boundary-boundary list to rows of K slots (it is read once, when the walls move),
//     the per-particle count to cnt[i]; overflow past the capacity is counted and reported
//     so the host can grow the capacity and search again;
//   * the fluid search stages a workgroup's candidates in LDS, row by row, so that each
//     particle walks its 27 cells as 9 contiguous runs (k_neighbors_fluid).
#include "device_common.hpp"
#include "kernels.hpp"

#pragma clang fp contract(off)

namespace mps {
namespace dev {

// ===================================================================== radix sort
constexpr int kSortThreads = 256;
constexpr int kSortRounds = 8;
constexpr int kSortTile = kSortThreads * kSortRounds;

__global__ void __launch_bounds__(kSortThreads) k_radix_hist(const uint32_t* __restrict__ keys, uint32_t n, int shift,
                                                           uint32_t* __restrict__ hist, uint32_t nblk) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0u;
    __syncthreads();
    const uint32_t base = blockIdx.x * kSortTile;
#pragma unroll
    for (int r = 0; r < kSortRounds; r++) {
        const uint32_t i = base + r * kSortThreads + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * nblk + blockIdx.x] = h[threadIdx.x];
}

// Exclusive scan of the digit histogram (length L = 256 * tiles), reduce-then-scan over
// 1024-element chunks: chunk sums, one-block scan of the sums, chunk scans.
constexpr int kScanChunk = 1024;
__device__ __forceinline__ uint32_t block_incl_scan_1024(uint32_t v, uint32_t* lds /*[16]*/, uint32_t* total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t t = __shfl_up(v, off, 64);
        if (lane >= off) v += t;
    }
    if (lane == 63) lds[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int q = 0; q < (int)(blockDim.x >> 6); q++) {
            const uint32_t t = lds[q];
            lds[q] = run;
            run += t;
        }
        lds[16] = run;
    }
    __syncthreads();
    v += lds[w];
    *total = lds[16];
    __syncthreads();
    return v;
}
__global__ void __launch_bounds__(kScanChunk) k_scan_sums(const uint32_t* __restrict__ a, uint32_t L, uint32_t* __restrict__ sums) {
    __shared__ uint32_t lds[17];
    const uint32_t i = blockIdx.x * kScanChunk + threadIdx.x;
    uint32_t tot;
    block_incl_scan_1024(i < L ? a[i] : 0u, lds, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}
__global__ void __launch_bounds__(kScanChunk) k_scan_top(uint32_t* __restrict__ sums, uint32_t n) {
    __shared__ uint32_t lds[17];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < n; base += kScanChunk) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t v = i < n ? sums[i] : 0u;
        uint32_t tot;
        const uint32_t inc = block_incl_scan_1024(v, lds, &tot);
        if (i < n) sums[i] = carry + inc - v;
        carry += tot;
    }
}
__global__ void __launch_bounds__(kScanChunk) k_scan_apply(uint32_t* __restrict__ a, uint32_t L, const uint32_t* __restrict__ sums) {
    __shared__ uint32_t lds[17];
    const uint32_t i = blockIdx.x * kScanChunk + threadIdx.x;
    const uint32_t v = i < L ? a[i] : 0u;
    uint32_t tot;
    const uint32_t inc = block_incl_scan_1024(v, lds, &tot);
    if (i < L) a[i] = sums[blockIdx.x] + inc - v;
}

// Stable scatter: items are ranked round by round (256 per round, wave-major, lane-minor),
// which preserves input order among equal digits.
__global__ void __launch_bounds__(kSortThreads) k_radix_scatter(const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                              uint32_t* __restrict__ kout, uint32_t* __restrict__ vout, uint32_t n,
                                                              int shift, const uint32_t* __restrict__ offs, uint32_t nblk) {
    __shared__ uint32_t base[256];
    __shared__ uint32_t wcnt[4][256];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    base[threadIdx.x] = offs[threadIdx.x * nblk + blockIdx.x];
    const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
    for (int r = 0; r < kSortRounds; r++) {
#pragma unroll
        for (int k = 0; k < 4; k++) wcnt[k][threadIdx.x] = 0u;
        __syncthreads();
        const uint32_t i = blockIdx.x * kSortTile + r * kSortThreads + threadIdx.x;
        const bool valid = i < n;
        const uint32_t key = valid ? kin[i] : 0u;
        const uint32_t d = (key >> shift) & 255u;
        uint64_t m = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bb = __ballot(bit);
            m &= bit ? bb : ~bb;
        }
        const uint32_t rank = __popcll(m & lt);
        if (valid && rank == 0u) wcnt[w][d] = __popcll(m);
        __syncthreads();
        {
            const uint32_t t = threadIdx.x;
            const uint32_t c0 = wcnt[0][t], c1 = wcnt[1][t], c2 = wcnt[2][t], c3 = wcnt[3][t];
            const uint32_t b0 = base[t];
            wcnt[0][t] = b0;
            wcnt[1][t] = b0 + c0;
            wcnt[2][t] = b0 + c0 + c1;
            wcnt[3][t] = b0 + c0 + c1 + c2;
            base[t] = b0 + c0 + c1 + c2 + c3;
        }
        __syncthreads();
        if (valid) {
            const uint32_t pos = wcnt[w][d] + rank;
            kout[pos] = key;
            vout[pos] = vin[i];
        }
        __syncthreads();
    }
}

// ===================================================================== hashing
// Keys of a set held in id order. With several scenes the scene index sits above the cell
// bits (scene << shift), so a sort keeps every scene contiguous; `scene` is null for one.
__global__ void __launch_bounds__(kBlock) k_cell_keys(const float4* __restrict__ pos, uint32_t n, HashGeom g, int morton,
                                                    const uint32_t* __restrict__ scene, int shift,
                                                    uint32_t* __restrict__ key, uint32_t* __restrict__ id) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 p = pos[i];
    const i3 c = grid_pos(p.x, p.y, p.z, g);
    const uint32_t k = morton ? grid_zindex(c, g) : grid_index(c.x, c.y, c.z, g);
    key[i] = scene ? (scene[i] << shift) | k : k;
    id[i] = i;
}

// Keys of particles held in slot order, written at their reference index: keyref[perm[s]]
// (scene index above the cell bits). Padding slots between scenes (perm = kEmpty) are skipped.
__global__ void __launch_bounds__(kBlock) k_cell_keys_slots(const float4* __restrict__ pos, const uint32_t* __restrict__ perm,
                                                          uint32_t n, HashGeom g, int morton, const uint32_t* __restrict__ blk_scene,
                                                          int shift, uint32_t* __restrict__ keyref, uint32_t* __restrict__ iota) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    iota[s] = s;
    const uint32_t ref = perm[s];
    if (ref == kEmpty) return;
    const float4 p = pos[s];
    const i3 c = grid_pos(p.x, p.y, p.z, g);
    const uint32_t k = morton ? grid_zindex(c, g) : grid_index(c.x, c.y, c.z, g);
    keyref[ref] = (blk_scene[blockIdx.x] << shift) | k;
}
// Morton code of the (wrapped) cell of each slot, written at the slot's reference index;
// sorting by it keeps each cell contiguous and blocks of slots spatially compact.
__device__ __forceinline__ uint32_t cell_morton(i3 c, const HashGeom& g) {
    const uint32_t x = static_cast<uint32_t>(c.x) & (g.hx - 1u);
    const uint32_t y = static_cast<uint32_t>(c.y) & (g.hy - 1u);
    const uint32_t z = static_cast<uint32_t>(c.z) & (g.hz - 1u);
    return split_by3(x) | (split_by3(y) << 1u) | (split_by3(z) << 2u);
}
__global__ void __launch_bounds__(kBlock) k_cell_morton_slots(const float4* __restrict__ pos, const uint32_t* __restrict__ perm,
                                                            uint32_t n, HashGeom g, const uint32_t* __restrict__ blk_scene,
                                                            int shift, uint32_t* __restrict__ keyref, uint32_t* __restrict__ iota) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    iota[s] = s;
    const uint32_t ref = perm[s];
    if (ref == kEmpty) return;
    const float4 p = pos[s];
    keyref[ref] = (blk_scene[blockIdx.x] << shift) | cell_morton(grid_pos(p.x, p.y, p.z, g), g);
}
// Sorted position q (key[q], reference id id[q]) -> slot q + delta[scene]: scene s's sorted
// range [cum_s, cum_s + count_s) lands on its slots [begin_s, end_s). Writes the new perm, the
// old slot of each particle (src, for the gathers) and the key per slot.
__global__ void __launch_bounds__(kBlock) k_place(const uint32_t* __restrict__ key, const uint32_t* __restrict__ id, uint32_t nref,
                                                int shift, const uint32_t* __restrict__ delta, const uint32_t* __restrict__ inv,
                                                uint32_t* __restrict__ perm, uint32_t* __restrict__ src,
                                                uint32_t* __restrict__ mkey) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nref) return;
    const uint32_t k = key[q];
    const uint32_t slot = q + delta[k >> shift];
    const uint32_t ref = id[q];
    perm[slot] = ref;
    src[slot] = inv[ref];
    mkey[slot] = k;
}
// Cell table over Morton-sorted slots: start/end are indexed by scene * C + the reference's
// linear cell key; lkey[s] keeps that index (for clearing and for exporting the hash). mkey
// is kEmpty on padding slots, which therefore end every run.
__global__ void __launch_bounds__(kBlock) k_cell_table_m(const uint32_t* __restrict__ mkey, const float4* __restrict__ pos,
                                                       uint32_t n, HashGeom g, int shift, uint32_t C,
                                                       uint32_t* __restrict__ start, uint32_t* __restrict__ end,
                                                       uint32_t* __restrict__ lkey) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const uint32_t m = mkey[s];
    if (m == kEmpty) return;
    const float4 p = pos[s];
    const i3 c = grid_pos(p.x, p.y, p.z, g);
    const uint32_t t = (m >> shift) * C + grid_index(c.x, c.y, c.z, g);
    lkey[s] = t;
    if (s == 0 || mkey[s - 1] != m) start[t] = s;
    if (s == n - 1 || mkey[s + 1] != m) end[t] = s + 1;
}
__global__ void __launch_bounds__(kBlock) k_clear_cells_m(const uint32_t* __restrict__ mkey, const uint32_t* __restrict__ lkey,
                                                        uint32_t n, uint32_t* __restrict__ start, uint32_t* __restrict__ end) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const uint32_t m = mkey[s];
    if (m == kEmpty) return;
    if (s == 0 || mkey[s - 1] != m) {
        start[lkey[s]] = kEmpty;
        end[lkey[s]] = kEmpty;
    }
}
// inv[perm[s]] = s (padding slots skipped)
__global__ void __launch_bounds__(kBlock) k_invert(const uint32_t* __restrict__ perm, uint32_t n, uint32_t* __restrict__ inv) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n && perm[s] != kEmpty) inv[perm[s]] = s;
}
// src[s'] = inv[perm_new[s']]: old slot of the particle that moves to slot s'
__global__ void __launch_bounds__(kBlock) k_compose(const uint32_t* __restrict__ perm_new, const uint32_t* __restrict__ inv,
                                                  uint32_t n, uint32_t* __restrict__ src) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) src[s] = inv[perm_new[s]];
}
// perm[s] = map[perm[s]]  (relabel reference ids after a Z-sort)
__global__ void __launch_bounds__(kBlock) k_relabel(uint32_t* __restrict__ perm, const uint32_t* __restrict__ map, uint32_t n) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n && perm[s] != kEmpty) perm[s] = map[perm[s]];
}
__global__ void __launch_bounds__(kBlock) k_iota(uint32_t* __restrict__ p, uint32_t n) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) p[s] = s;
}

__global__ void __launch_bounds__(kBlock) k_fill_u32(uint32_t* __restrict__ p, size_t n, uint32_t v) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Reset only the cells that the previous sorted keys occupied.
__global__ void __launch_bounds__(kBlock) k_clear_cells(const uint32_t* __restrict__ key, uint32_t n, uint32_t* __restrict__ start,
                                                      uint32_t* __restrict__ end) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const uint32_t k = key[s];
    if (s == 0 || key[s - 1] != k) {
        start[k] = kEmpty;
        end[k] = kEmpty;
    }
}

// ReorderHash_kernel (MPSSpatialHash.cuh:115-145) + positions in key order.
__global__ void __launch_bounds__(kBlock) k_cell_table(const uint32_t* __restrict__ key, const uint32_t* __restrict__ id, uint32_t n,
                                                     uint32_t* __restrict__ start, uint32_t* __restrict__ end,
                                                     const float4* __restrict__ pos, float4* __restrict__ spos) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const uint32_t k = key[s];
    if (s == 0 || key[s - 1] != k) {
        start[k] = s;
        if (s > 0) end[key[s - 1]] = s;
    }
    if (s == n - 1) end[k] = n;
    if (spos) spos[s] = pos[id[s]];
}

// 27-cell scan (Research, MPSSpatialHash.cuh:47-80): z outer, y, x inner; slots in key order.
// Neighbour when |x_i - x_j| < (4 r_i + 4 r_j) * 0.5, self excluded when SELF.
// The 27 (start, end) pairs are loaded up front, so their latencies overlap instead of
// adding up cell after cell; candidates are read two at a time for the same reason. Lists
// are written column-major (out[k*stride + i]) when COLS, else as rows of K slots.
// Neighbour test of the reference: |x_i - x_j| < (h_i + h_j) / 2 with h = 4 r.
// sqrt is monotone and correctly rounded: far from the cut-off the squared distance decides;
// near it the exact test of the reference runs.
__device__ __forceinline__ bool within(const float4& xi, float ri, const float4& q) {
    const float ddx = xi.x - q.x, ddy = xi.y - q.y, ddz = xi.z - q.z;
    const float dd = (ddx * ddx + ddy * ddy) + ddz * ddz;
    const float rj = q.w * 4.0f;
    const float thr = (ri + rj) * 0.5f;
    const float t2 = thr * thr;
    if (dd < t2 * 0.9999f) return true;
    if (dd > t2 * 1.0001f) return false;
    return __builtin_sqrtf(dd) < thr;
}

// within() for four candidates of one radius, as a bit mask (bit k: candidate k). The cut-off
// thr = (h_i + h_j) * 0.5 (h = 4 r) is then the same for every pair, and it, t2lo = thr^2 * 0.9999
// and t2hi = thr^2 * 1.0001 come precomputed with the same float operations. The squared distances decide first, and the square
// root is taken only within 1e-4 of the cut-off. The two range tests are sign bits rather than
// comparisons: for finite values dd - t2lo is negative exactly when dd < t2lo, and t2hi - dd
// exactly when dd > t2hi (two unequal floats differ by a nonzero float; equal ones give +0).
// A NaN distance sets both bits or neither, and is then no hit.
__device__ __forceinline__ uint32_t hit_bits4(const float4& xi, float thr, float t2lo, float t2hi, const float4 (&q)[4]) {
    uint32_t lo = 0u, hi = 0u;
    float dd[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const float ddx = xi.x - q[k].x, ddy = xi.y - q[k].y, ddz = xi.z - q[k].z;
        dd[k] = (ddx * ddx + ddy * ddy) + ddz * ddz;
        lo |= (__float_as_uint(dd[k] - t2lo) >> 31) << k;
        hi |= (__float_as_uint(t2hi - dd[k]) >> 31) << k;
    }
    uint32_t in = lo & ~hi;
    const uint32_t near = ~(lo | hi) & 15u;
    if (near) {
#pragma unroll
        for (int k = 0; k < 4; k++)
            if ((near >> k) & 1u) in |= static_cast<uint32_t>(__builtin_sqrtf(dd[k]) < thr) << k;
    }
    return in;
}

template <bool SELF, bool COLS>
__device__ __forceinline__ uint32_t search_cells(float4 xi, uint32_t i, const HashGeom& g, const uint32_t* __restrict__ start,
                                                 const uint32_t* __restrict__ end, const uint32_t* __restrict__ ids,
                                                 const float4* __restrict__ spos, uint32_t* __restrict__ out, uint32_t K,
                                                 uint32_t stride) {
    const i3 p = grid_pos(xi.x, xi.y, xi.z, g);
    const float ri = xi.w * 4.0f;
    uint32_t k = 0u;
    auto test = [&](float4 q, uint32_t s) {
        if (within(xi, ri, q)) {
            const uint32_t j = ids ? ids[s] : s;
            if (!(SELF && j == i)) {
                if (k < K) out[COLS ? size_t(k) * stride + i : size_t(i) * K + k] = j;
                k++;
            }
        }
    };
    // the 27 cells in the reference order; the loop stays rolled (an unrolled walk does not fit
    // the instruction cache)
#pragma unroll 1
    for (int t = 0; t < 27; t++) {
        const uint32_t c = grid_index(p.x + (t % 3) - 1, p.y + (t / 3) % 3 - 1, p.z + t / 9 - 1, g);
        uint32_t s = start[c];
        const uint32_t e = end[c];
        // an empty cell has start = end = 0xFFFFFFFF: only e - s (never s + 1) is formed,
        // and only once s < e is known
        if (s < e) {
            for (; e - s >= 2u; s += 2) {
                const float4 q0 = spos[s], q1 = spos[s + 1];
                test(q0, s);
                test(q1, s + 1);
            }
            if (s < e) test(spos[s], s);
        }
    }
    return k;
}

// Per-workgroup (largest count, sum of counts) of up to two lists, written to st[blk] and
// st[n + blk]; k_count_fold reduces them (no same-address atomics from every wave).
template <int L>
__device__ __forceinline__ void block_counts(const uint32_t (&c)[L], uint2* __restrict__ st, uint32_t n, uint32_t blk) {
    __shared__ uint32_t wm[L][kBlock / 64], ws[L][kBlock / 64];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
#pragma unroll
    for (int l = 0; l < L; l++) {
        uint32_t mx = c[l], sm = c[l];
        for (int off = 32; off > 0; off >>= 1) {
            mx = max(mx, (uint32_t)__shfl_xor((int)mx, off, 64));
            sm += (uint32_t)__shfl_xor((int)sm, off, 64);
        }
        if (lane == 0) {
            wm[l][wave] = mx;
            ws[l][wave] = sm;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int l = 0; l < L; l++) {
            uint32_t mx = 0u, sm = 0u;
            for (uint32_t w = 0; w < kBlock / 64; w++) {
                mx = max(mx, wm[l][w]);
                sm += ws[l][w];
            }
            st[size_t(l) * n + blk] = make_uint2(mx, sm);
        }
    }
}

constexpr int kCountFoldThreads = 1024;
__global__ void __launch_bounds__(kCountFoldThreads) k_count_fold(const uint2* __restrict__ st, uint32_t n, uint32_t* maxc,
                                                                unsigned long long* nnz) {
    __shared__ uint32_t wm[kCountFoldThreads / 64];
    __shared__ unsigned long long ws[kCountFoldThreads / 64];
    uint32_t mx = 0u;
    unsigned long long sm = 0ull;
    for (uint32_t k = threadIdx.x; k < n; k += blockDim.x) {
        const uint2 v = st[k];
        mx = max(mx, v.x);
        sm += v.y;
    }
    for (int off = 32; off > 0; off >>= 1) {
        mx = max(mx, (uint32_t)__shfl_xor((int)mx, off, 64));
        sm += __shfl_xor(sm, off, 64);
    }
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    if (lane == 0) {
        wm[wave] = mx;
        ws[wave] = sm;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (uint32_t w = 1; w < blockDim.x / 64; w++) {
            mx = max(mx, wm[w]);
            sm += ws[w];
        }
        *maxc = mx;
        *nnz = sm;
    }
}

// Scene s's cells are entries [s*C, (s+1)*C) of both tables.
//
// A workgroup owns 256 consecutive fluid slots (sorted by cell, so each cell's particles are
// one run). The first particle of each run registers the 9 (y, z) rows of its 27 cells in an
// LDS hash of the block's rows (a row keyed by its wrapped y, z), widening the row's x range to
// cover x-1..x+1. Each row's cells are then numbered in x order over that range, each cell's
// bounds are read once, and all fluid (then wall) particles of those cells are staged in LDS
// in that order, each with its slot (a wall particle: its id) in place of the radius. So the
// three cells x-1, x, x+1 of any row of a particle are one contiguous run of staged copies, and
// every particle walks its 27 cells in the reference order as 9 row runs. A block with more
// rows, cells or candidates than the LDS holds, or whose staged fluid (or wall) particles do
// not all share one radius, searches the global tables for that set; the lists are the same
// either way.
constexpr uint32_t kSrchRowHash = 128;  // LDS hash slots for rows
constexpr uint32_t kSrchRows = 96;      // rows per block
constexpr uint32_t kSrchCells = 384;    // cells per block (the rows' x ranges)
constexpr uint32_t kSrchCand = 1920;    // staged candidates (fluid, then walls)
constexpr uint32_t kSrchPad = 4;        // a group of four read at the end of a run stays in the array
constexpr uint32_t kNoRow = 0xFFFFFFFFu;
__device__ __forceinline__ uint32_t row_hash(uint32_t c) { return (c * 2654435761u) >> 25; }
static_assert(kSrchRowHash == (1u << 7), "row_hash gives 7 bits");
static_assert(kSrchRows <= kBlock && kSrchCells <= 2u * kBlock, "one row per thread, two cells per thread");

// Inclusive scan of a 64-bit value over the kBlock threads; `ws` holds kBlock / 64 wave totals
// (synchronise before reusing it). `total` receives the block's sum.
__device__ __forceinline__ unsigned long long block_scan_u64(unsigned long long v, unsigned long long* ws,
                                                             unsigned long long& total) {
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    for (uint32_t off = 1; off < 64u; off <<= 1) {
        const uint32_t lo = __shfl_up(static_cast<uint32_t>(v), off, 64);
        const uint32_t hi = __shfl_up(static_cast<uint32_t>(v >> 32), off, 64);
        if (lane >= off) v += (static_cast<unsigned long long>(hi) << 32) | lo;
    }
    if (lane == 63u) ws[w] = v;
    __syncthreads();
    total = 0ull;
    for (uint32_t q = 0; q < kBlock / 64u; q++) {
        if (q < w) v += ws[q];
        total += ws[q];
    }
    return v;
}

// The hits among staged candidates [lo, hi) in order: four candidates per step (hit_bits4,
// the threshold of equal radii), a hit mask per 32 candidates, then the hits' ids (held in w) as
// list entries cnt, cnt + 1, ... (column-major, stride N; entries past K are counted only).
// `self` is the particle's own staged index, which is not a neighbour (any value >= hi: none).
__device__ __forceinline__ void walk_run(const float4* cand, uint32_t lo, uint32_t hi, const float4& xi, float thr, float t2lo,
                                         float t2hi, uint32_t self, uint32_t* __restrict__ out, uint32_t K, uint32_t N,
                                         uint32_t i, uint32_t& cnt) {
    for (uint32_t c0 = lo; c0 < hi; c0 += 32u) {
        const uint32_t n = min(hi - c0, 32u);
        uint32_t m = 0u;
        for (uint32_t q = 0; q < n; q += 4u) {
            float4 v[4];
#pragma unroll
            for (uint32_t k = 0; k < 4u; k++) v[k] = cand[c0 + q + k];  // <= hi + 2 < array size
            m |= hit_bits4(xi, thr, t2lo, t2hi, v) << q;
        }
        m &= n < 32u ? (1u << n) - 1u : ~0u;  // the group of four read past hi
        const uint32_t d = self - c0;
        if (d < 32u) m &= ~(1u << d);
        while (m) {
            const uint32_t b = static_cast<uint32_t>(__builtin_ctz(m));
            m &= m - 1u;
            if (cnt < K) out[size_t(cnt) * N + i] = __float_as_uint(cand[c0 + b].w);
            cnt++;
        }
    }
}

__global__ void __launch_bounds__(kBlock) k_neighbors_fluid(const float4* __restrict__ pos, Scenes sc, uint32_t C, HashGeom g,
                                                          HashView fl, HashView bd, uint32_t* __restrict__ nff, uint32_t Kff,
                                                          uint32_t* __restrict__ cff, uint32_t* __restrict__ nfb, uint32_t Kfb,
                                                          uint32_t* __restrict__ cfb, uint2* __restrict__ bstat) {
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t tid = threadIdx.x;
    const uint32_t i = blk * kBlock + tid;
    const uint32_t N = sc.N;
    const uint32_t sid = sc.blk_scene[blk];
    const uint32_t end = sc.P[sid].end;
    const bool live = i < end;
    const size_t base = size_t(sid) * C;
    const uint32_t* fstart = fl.start + base;
    const uint32_t* fend = fl.end + base;
    const uint32_t* bstart = bd.start + base;
    const uint32_t* bend = bd.end + base;
    __shared__ uint32_t rkey[kSrchRowHash];      // row hash: wrapped key of the row's cell x = 0
    __shared__ int rxlo[kSrchRowHash], rxhi[kSrchRowHash];  // the row's x range (unwrapped)
    __shared__ uint16_t rbase[kSrchRowHash];     // the row's first cell
    __shared__ uint8_t rslot[kSrchRows];         // hash slot of each row, in registration order
    __shared__ uint32_t ecell[kSrchCells], efs[kSrchCells], ebs[kSrchCells];
    __shared__ uint16_t efo[kSrchCells + 1], ebo[kSrchCells + 1];
    __shared__ float4 cand[kSrchCand + kSrchPad];
    __shared__ unsigned long long ws[kBlock / 64];
    __shared__ uint32_t nrow, mixed, first[2];  // first: first cell with fluid, with walls
    if (tid < kSrchRowHash) {
        rkey[tid] = kNoRow;
        rxlo[tid] = 0x7FFFFFFF;
        rxhi[tid] = -0x7FFFFFFF - 1;
    }
    if (tid == 0) {
        nrow = 0u;
        mixed = 0u;
        first[0] = first[1] = kNoRow;
    }
    const float4 xi = live ? pos[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    const i3 p = grid_pos(xi.x, xi.y, xi.z, g);
    bool leader = live;
    if (live && tid > 0) {
        const float4 xp = pos[i - 1];
        const i3 pp = grid_pos(xp.x, xp.y, xp.z, g);
        leader = pp.x != p.x || pp.y != p.y || pp.z != p.z;
    }
    __syncthreads();
    // 1. rows (inserts stop once the block is known to be over the cap, so the table never
    //    fills; every probe loop is bounded anyway)
    if (leader) {
        for (int r = 0; r < 9; r++) {
            if (*(volatile uint32_t*)&nrow > kSrchRows) break;
            const uint32_t k = grid_index(0, p.y + r % 3 - 1, p.z + r / 3 - 1, g);
            uint32_t h = row_hash(k);
            for (uint32_t probe = 0; probe < kSrchRowHash; probe++) {
                const uint32_t prev = atomicCAS(&rkey[h], kNoRow, k);
                if (prev == kNoRow) {
                    const uint32_t e = atomicAdd(&nrow, 1u);
                    if (e < kSrchRows) rslot[e] = static_cast<uint8_t>(h);
                }
                if (prev == kNoRow || prev == k) {
                    atomicMin(&rxlo[h], p.x - 1);
                    atomicMax(&rxhi[h], p.x + 1);
                    break;
                }
                h = (h + 1u) & (kSrchRowHash - 1u);
            }
        }
    }
    __syncthreads();
    const uint32_t nr = nrow;
    bool staged = nr <= kSrchRows;
    uint32_t nc = 0u;
    // 2. number each row's cells over its x range (rows in registration order)
    if (staged) {
        const uint32_t h = tid < nr ? rslot[tid] : 0u;
        const uint32_t len = tid < nr ? static_cast<uint32_t>(min(static_cast<long long>(rxhi[h]) - rxlo[h] + 1, 
                                                                 static_cast<long long>(kSrchCells) + 1)) : 0u;
        unsigned long long tot;
        const uint32_t b0 = static_cast<uint32_t>(block_scan_u64(len, ws, tot)) - len;
        nc = static_cast<uint32_t>(tot);
        staged = nc <= kSrchCells;
        if (staged && tid < nr) {
            rbase[h] = static_cast<uint16_t>(b0);
            const uint32_t k0 = rkey[h];
            for (uint32_t q = 0; q < len; q++) ecell[b0 + q] = k0 + (static_cast<uint32_t>(rxlo[h] + static_cast<int>(q)) & (g.hx - 1u));
        }
        __syncthreads();
    }
    // 3. bounds of each cell (read once) and offsets of its staged copies: fluid from 0, walls
    //    after all fluid
    uint32_t tf = 0u, tb = 0u;
    if (staged) {
        uint32_t fc[2] = {0u, 0u}, bc[2] = {0u, 0u};
#pragma unroll
        for (uint32_t q = 0; q < 2; q++) {
            const uint32_t e = 2u * tid + q;
            if (e < nc) {
                const uint32_t c = ecell[e];
                const uint32_t s0 = fstart[c], s1 = fend[c];
                efs[e] = s0;
                fc[q] = s0 < s1 ? s1 - s0 : 0u;  // an empty cell has start = end = 0xFFFFFFFF
                if (bd.n) {
                    const uint32_t b0 = bstart[c], b1 = bend[c];
                    ebs[e] = b0;
                    bc[q] = b0 < b1 ? b1 - b0 : 0u;
                }
            }
        }
        const unsigned long long mine = (static_cast<unsigned long long>(bc[0] + bc[1]) << 32) | (fc[0] + fc[1]);
        unsigned long long tot;
        const unsigned long long ex = block_scan_u64(mine, ws, tot) - mine;
        tf = static_cast<uint32_t>(tot & 0xFFFFFFFFull);
        tb = static_cast<uint32_t>(tot >> 32);
        staged = static_cast<unsigned long long>(tf) + tb <= kSrchCand;
        uint32_t fo = static_cast<uint32_t>(ex & 0xFFFFFFFFull), bo = static_cast<uint32_t>(ex >> 32);
        if (staged) {
            // each copy's w first holds its cell (read back by the staging pass)
#pragma unroll
            for (uint32_t q = 0; q < 2; q++) {
                const uint32_t e = 2u * tid + q;
                if (e < nc) {
                    efo[e] = static_cast<uint16_t>(fo);
                    ebo[e] = static_cast<uint16_t>(tf + bo);
                    for (uint32_t k = 0; k < fc[q]; k++) cand[fo + k].w = __uint_as_float(e);
                    for (uint32_t k = 0; k < bc[q]; k++) cand[tf + bo + k].w = __uint_as_float(e);
                    if (fc[q]) atomicMin(&first[0], e);
                    if (bc[q]) atomicMin(&first[1], e);
                }
                fo += fc[q];
                bo += bc[q];
            }
            if (tid == 0) {
                efo[nc] = static_cast<uint16_t>(tf);
                ebo[nc] = static_cast<uint16_t>(tf + tb);
            }
        }
        __syncthreads();
    }
    // 4. stage the candidates with their ids (copy t of cell e is particle efs[e] + t - efo[e]).
    //    `mixed` records whether some fluid (bit 0) or wall (bit 1) candidate has another radius
    //    than the first staged candidate of its set.
    float rw = 0.0f;  // the walls' common radius
    if (staged) {
        uint32_t mx = 0u;
        if (tf) {
            const float rf = pos[efs[first[0]]].w;
            for (uint32_t t = tid; t < tf; t += kBlock) {
                const uint32_t e = __float_as_uint(cand[t].w);
                const uint32_t s = efs[e] + (t - efo[e]);
                const float4 v = pos[s];
                cand[t] = make_float4(v.x, v.y, v.z, __uint_as_float(s));
                mx |= v.w != rf ? 1u : 0u;
            }
        }
        if (tb) {
            rw = bd.spos[ebs[first[1]]].w;
            for (uint32_t t = tf + tid; t < tf + tb; t += kBlock) {
                const uint32_t e = __float_as_uint(cand[t].w);
                const uint32_t sb = ebs[e] + (t - ebo[e]);
                const float4 v = bd.spos[sb];
                cand[t] = make_float4(v.x, v.y, v.z, __uint_as_float(bd.id ? bd.id[sb] : sb));
                mx |= v.w != rw ? 2u : 0u;
            }
        }
        if (mx) atomicOr(&mixed, mx);
        __syncthreads();
    }
    // 5. the search proper: 9 rows (z, then y offsets), each the run of cells x-1, x, x+1
    uint32_t a = 0u, b = 0u;
    if (live) {
        const uint32_t mix = staged ? mixed : 3u;
        const float ri = xi.w * 4.0f;
        // the particle's cell x + dx in row (y, z)
        auto cell_of = [&](int dx, int y, int z) -> uint32_t {
            const uint32_t k = grid_index(0, y, z, g);
            uint32_t h = row_hash(k);
#pragma unroll 1
            for (uint32_t probe = 0; probe < kSrchRowHash && rkey[h] != k; probe++) h = (h + 1u) & (kSrchRowHash - 1u);
            return rkey[h] == k ? rbase[h] + static_cast<uint32_t>(p.x + dx - rxlo[h]) : kNoRow;
        };
        auto rows = [&](const uint16_t* eo, float thr, uint32_t self, uint32_t* out, uint32_t K, uint32_t& cnt) {
            const float t2 = thr * thr, t2lo = t2 * 0.9999f, t2hi = t2 * 1.0001f;
            for (int r = 0; r < 9; r++) {
                const uint32_t e = cell_of(-1, p.y + r % 3 - 1, p.z + r / 3 - 1);
                if (e == kNoRow) continue;  // not reached: every row of a run was registered
                walk_run(cand, eo[e], eo[e + 3u], xi, thr, t2lo, t2hi, self, out, K, N, i, cnt);
            }
        };
        if (!(mix & 1u)) {
            const uint32_t eh = cell_of(0, p.y, p.z);
            const uint32_t self = eh == kNoRow ? kNoRow : efo[eh] + (i - efs[eh]);
            rows(efo, (ri + ri) * 0.5f, self, nff, Kff, a);
        } else {
            a = search_cells<true, true>(xi, i, g, fstart, fend, fl.id, fl.spos, nff, Kff, N);
        }
        if (bd.n) {
            if (staged && tb == 0u) {
                b = 0u;  // no wall particle in any cell of the block
            } else if (!(mix & 2u)) {
                rows(ebo, (ri + rw * 4.0f) * 0.5f, kNoRow, nfb, Kfb, b);
            } else {
                b = search_cells<false, true>(xi, i, g, bstart, bend, bd.id, bd.spos, nfb, Kfb, N);
            }
        }
    }
    if (i < N) {
        cff[i] = a;
        cfb[i] = b;
    }
    const uint32_t cnt[2] = {a, b};
    block_counts<2>(cnt, bstat, gridDim.x, blk);
}

__global__ void __launch_bounds__(kBlock) k_neighbors_boundary(const float4* __restrict__ pos, uint32_t B,
                                                             const uint32_t* __restrict__ scene, uint32_t C, HashGeom g,
                                                             HashView bd, uint32_t* __restrict__ nbb, uint32_t Kbb,
                                                             uint32_t* __restrict__ cbb, uint2* __restrict__ bstat) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t a = 0u;
    if (i < B) {
        const size_t base = scene ? size_t(scene[i]) * C : 0;
        a = search_cells<true, false>(pos[i], i, g, bd.start + base, bd.end + base, bd.id, bd.spos, nbb, Kbb, B);
        cbb[i] = a;
    }
    const uint32_t cnt[1] = {a};
    block_counts<1>(cnt, bstat, gridDim.x, blockIdx.x);
}

// Ref lists of coupled fluid objects (SpatialHash::BuildNeighorhood(obj, ref, refHash),
// MPSSpatialHash.cu:184-218, .cuh:180-271): for each fluid slot, the particles of every other
// scene within (h_i + h_j)/2, scene after scene in ascending order, each scene's part in the
// 27-cell order of that scene's own table (the walk of the wall lists, on another table).
__global__ void __launch_bounds__(kBlock) k_neighbors_ref(const float4* __restrict__ pos, Scenes sc, uint32_t C, HashGeom g,
                                                        HashView fl, uint32_t* __restrict__ nfr, uint32_t Kfr,
                                                        uint32_t* __restrict__ cfr, uint2* __restrict__ bstat) {
    const uint32_t blk = blockIdx.x;
    const uint32_t i = blk * kBlock + threadIdx.x;
    const uint32_t N = sc.N;
    const uint32_t sid = sc.blk_scene[blk];
    uint32_t k = 0u;
    if (i < sc.P[sid].end) {
        const float4 xi = pos[i];
        for (uint32_t t = 0; t < sc.S; t++) {
            if (t == sid) continue;
            const uint32_t at = k < Kfr ? k : Kfr;  // entries past the capacity are counted, not written
            k += search_cells<false, true>(xi, i, g, fl.start + size_t(t) * C, fl.end + size_t(t) * C, nullptr, fl.spos,
                                           nfr + size_t(at) * N, Kfr - at, N);
        }
    }
    if (i < N) cfr[i] = k;
    const uint32_t cnt[1] = {k};
    block_counts<1>(cnt, bstat, gridDim.x, blk);
}

// Z-sort payload permutation (MPSSpatialHash.cu:101-140): out[s] = in[perm[s]].
__global__ void __launch_bounds__(kBlock) k_gather_f4(const uint32_t* __restrict__ perm, uint32_t n, const float4* __restrict__ in,
                                                    float4* __restrict__ out) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n && perm[s] != kEmpty) out[s] = in[perm[s]];
}
__global__ void __launch_bounds__(kBlock) k_gather_u32(const uint32_t* __restrict__ perm, uint32_t n, const uint32_t* __restrict__ in,
                                                     uint32_t* __restrict__ out) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n && perm[s] != kEmpty) out[s] = in[perm[s]];
}
__global__ void __launch_bounds__(kBlock) k_gather_f2(const uint32_t* __restrict__ perm, uint32_t n, const float2* __restrict__ in,
                                                    float2* __restrict__ out) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n && perm[s] != kEmpty) out[s] = in[perm[s]];
}
__global__ void __launch_bounds__(kBlock) k_gather_f1(const uint32_t* __restrict__ perm, uint32_t n, const float* __restrict__ in,
                                                    float* __restrict__ out) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n && perm[s] != kEmpty) out[s] = in[perm[s]];
}

}  // namespace dev

// ===================================================================== host launchers
namespace launch {
using namespace dev;

static inline unsigned nb(size_t n, unsigned b = kBlock) { return unsigned((n + b - 1) / b); }

size_t radix_temp_words(uint32_t n) {
    const size_t L = size_t(256) * ((n + kSortTile - 1) / kSortTile);
    return L + (L + kScanChunk - 1) / kScanChunk + 1;
}

int radix_sort(uint32_t*& keys, uint32_t*& vals, uint32_t*& keys_alt, uint32_t*& vals_alt, uint32_t* temp, uint32_t n,
               int key_bits, hipStream_t s) {
    if (n == 0) return 0;
    const uint32_t nblk = (n + kSortTile - 1) / kSortTile;
    const uint32_t L = 256u * nblk;
    for (int shift = 0; shift < key_bits; shift += 8) {
        launch_k(k_radix_hist, nblk, kSortThreads, 0, s, keys, n, shift, temp, nblk);
        const uint32_t nch = (L + kScanChunk - 1) / kScanChunk;
        uint32_t* sums = temp + L;
        launch_k(k_scan_sums, nch, kScanChunk, 0, s, temp, L, sums);
        launch_k(k_scan_top, 1, kScanChunk, 0, s, sums, nch);
        launch_k(k_scan_apply, nch, kScanChunk, 0, s, temp, L, sums);
        launch_k(k_radix_scatter, nblk, kSortThreads, 0, s, keys, vals, keys_alt, vals_alt, n, shift, temp, nblk);
        std::swap(keys, keys_alt);
        std::swap(vals, vals_alt);
    }
    return 0;
}

void cell_keys(const float4* pos, uint32_t n, const HashGeom& g, int morton, const uint32_t* scene, int shift, uint32_t* key,
               uint32_t* id, hipStream_t s) {
    if (n) launch_k(k_cell_keys, nb(n), kBlock, 0, s, pos, n, g, morton, scene, shift, key, id);
}
void cell_keys_slots(const float4* pos, const uint32_t* perm, uint32_t n, const HashGeom& g, int morton,
                     const uint32_t* blk_scene, int shift, uint32_t* keyref, uint32_t* iota, hipStream_t s) {
    if (n) launch_k(k_cell_keys_slots, nb(n), kBlock, 0, s, pos, perm, n, g, morton, blk_scene, shift, keyref, iota);
}
void cell_morton_slots(const float4* pos, const uint32_t* perm, uint32_t n, const HashGeom& g, const uint32_t* blk_scene,
                       int shift, uint32_t* keyref, uint32_t* iota, hipStream_t s) {
    if (n) launch_k(k_cell_morton_slots, nb(n), kBlock, 0, s, pos, perm, n, g, blk_scene, shift, keyref, iota);
}
void place(const uint32_t* key, const uint32_t* id, uint32_t nref, int shift, const uint32_t* delta, const uint32_t* inv,
           uint32_t* perm, uint32_t* src, uint32_t* mkey, hipStream_t s) {
    if (nref) launch_k(k_place, nb(nref), kBlock, 0, s, key, id, nref, shift, delta, inv, perm, src, mkey);
}
void cell_table_m(const uint32_t* mkey, const float4* pos, uint32_t n, const HashGeom& g, int shift, uint32_t C, uint32_t* start,
                  uint32_t* end, uint32_t* lkey, hipStream_t s) {
    if (n) launch_k(k_cell_table_m, nb(n), kBlock, 0, s, mkey, pos, n, g, shift, C, start, end, lkey);
}
void clear_cells_m(const uint32_t* mkey, const uint32_t* lkey, uint32_t n, uint32_t* start, uint32_t* end, hipStream_t s) {
    if (n) launch_k(k_clear_cells_m, nb(n), kBlock, 0, s, mkey, lkey, n, start, end);
}
void invert(const uint32_t* perm, uint32_t n, uint32_t* inv, hipStream_t s) {
    if (n) launch_k(k_invert, nb(n), kBlock, 0, s, perm, n, inv);
}
void compose(const uint32_t* perm_new, const uint32_t* inv, uint32_t n, uint32_t* src, hipStream_t s) {
    if (n) launch_k(k_compose, nb(n), kBlock, 0, s, perm_new, inv, n, src);
}
void relabel(uint32_t* perm, const uint32_t* map, uint32_t n, hipStream_t s) {
    if (n) launch_k(k_relabel, nb(n), kBlock, 0, s, perm, map, n);
}
void iota(uint32_t* p, uint32_t n, hipStream_t s) {
    if (n) launch_k(k_iota, nb(n), kBlock, 0, s, p, n);
}
void fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t s) {
    if (n) launch_k(k_fill_u32, unsigned((n + kBlock - 1) / kBlock), kBlock, 0, s, p, n, v);
}
void clear_cells(const uint32_t* key, uint32_t n, uint32_t* start, uint32_t* end, hipStream_t s) {
    if (n) launch_k(k_clear_cells, nb(n), kBlock, 0, s, key, n, start, end);
}
void cell_table(const uint32_t* key, const uint32_t* id, uint32_t n, uint32_t* start, uint32_t* end, const float4* pos,
                float4* spos, hipStream_t s) {
    if (n) launch_k(k_cell_table, nb(n), kBlock, 0, s, key, id, n, start, end, pos, spos);
}
void neighbors_fluid(const float4* pos, const Scenes& sc, uint32_t C, const HashGeom& g, const HashView& fl, const HashView& bd,
                     uint32_t* nff, uint32_t Kff, uint32_t* cff, uint32_t* nfb, uint32_t Kfb, uint32_t* cfb, uint2* bstat,
                     NbrCounters* ctr, hipStream_t s) {
    if (!sc.N) return;
    const uint32_t g_ = nb(sc.N);
    launch_k(k_neighbors_fluid, g_, kBlock, 0, s, pos, sc, C, g, fl, bd, nff, Kff, cff, nfb, Kfb, cfb, bstat);
    launch_k(k_count_fold, 1, kCountFoldThreads, 0, s, bstat, g_, &ctr->max_ff, &ctr->nnz_ff);
    launch_k(k_count_fold, 1, kCountFoldThreads, 0, s, bstat + g_, g_, &ctr->max_fb, &ctr->nnz_fb);
}
void neighbors_ref(const float4* pos, const Scenes& sc, uint32_t C, const HashGeom& g, const HashView& fl, uint32_t* nfr,
                   uint32_t Kfr, uint32_t* cfr, uint2* bstat, NbrCounters* ctr, hipStream_t s) {
    if (!sc.N) return;
    const uint32_t g_ = nb(sc.N);
    launch_k(k_neighbors_ref, g_, kBlock, 0, s, pos, sc, C, g, fl, nfr, Kfr, cfr, bstat);
    launch_k(k_count_fold, 1, kCountFoldThreads, 0, s, bstat, g_, &ctr->max_fr, &ctr->nnz_fr);
}
void neighbors_boundary(const float4* pos, uint32_t B, const uint32_t* scene, uint32_t C, const HashGeom& g, const HashView& bd,
                        uint32_t* nbb, uint32_t Kbb, uint32_t* cbb, uint2* bstat, NbrCounters* ctr, hipStream_t s) {
    if (!B) return;
    const uint32_t g_ = nb(B);
    launch_k(k_neighbors_boundary, g_, kBlock, 0, s, pos, B, scene, C, g, bd, nbb, Kbb, cbb, bstat);
    launch_k(k_count_fold, 1, kCountFoldThreads, 0, s, bstat, g_, &ctr->max_bb, &ctr->nnz_bb);
}
void gather_f4(const uint32_t* perm, uint32_t n, const float4* in, float4* out, hipStream_t s) {
    if (n) launch_k(k_gather_f4, nb(n), kBlock, 0, s, perm, n, in, out);
}
void gather_u32(const uint32_t* perm, uint32_t n, const uint32_t* in, uint32_t* out, hipStream_t s) {
    if (n) launch_k(k_gather_u32, nb(n), kBlock, 0, s, perm, n, in, out);
}
void gather_f2(const uint32_t* perm, uint32_t n, const float2* in, float2* out, hipStream_t s) {
    if (n) launch_k(k_gather_f2, nb(n), kBlock, 0, s, perm, n, in, out);
}
void gather_f1(const uint32_t* perm, uint32_t n, const float* in, float* out, hipStream_t s) {
    if (n) launch_k(k_gather_f1, nb(n), kBlock, 0, s, perm, n, in, out);
}

}  // namespace launch
}  // namespace mps
