// Block halos of the fluid-fluid neighbour lists.
//
// The implicit viscosity operator (A.p) is applied ~30 times per step over the same lists.
// Gathering the 32-byte CG record of every neighbour from global memory costs most of its
// time: a block of 256 slots touches ~7k list entries but only ~900 distinct records (its
// halo), and the caches cannot keep the halos of all resident blocks. This pass, run once
// per solve, lists for each 256-slot block the distinct neighbour slots in ascending order
// (at most kHaloCap) and rewrites every list entry of the block as a 16-bit position in that
// list. A.p then stages the halo's records into LDS with coalesced loads and gathers from
// LDS. A block with more than kHaloCap distinct neighbours keeps the global path
// (halo count > kHaloCap marks it).
#include "device_common.hpp"
#include "kernels.hpp"

namespace mps {
namespace dev {

constexpr uint32_t kHaloHashBits = 11;
constexpr uint32_t kHaloHash = 1u << kHaloHashBits;  // LDS hash slots; occupancy stays below kHaloCap + kBlock
constexpr uint32_t kHaloSort = 2048;           // power of two >= kHaloCap
constexpr uint32_t kNoKey = 0xFFFFFFFFu;
static_assert(kHaloCap <= kHaloSort && kHaloCap + kBlock < kHaloHash, "halo table sizes");
static_assert(kHaloCap <= 65535u, "local positions are 16-bit");

__device__ __forceinline__ uint32_t halo_hash(uint32_t j) { return (j * 2654435761u) >> (32u - kHaloHashBits); }

// Slot of key j in the table (j must be present).
__device__ __forceinline__ uint32_t halo_find(const uint32_t* key, uint32_t j) {
    uint32_t s = halo_hash(j);
    for (uint32_t probe = 0; probe < kHaloHash; probe++) {
        if (key[s] == j) break;
        s = (s + 1u) & (kHaloHash - 1u);
    }
    return s;
}

// Blocks with more than `limit` (<= kHaloCap) distinct neighbours are marked not staged
// (n = kHaloCap + 1).
__global__ void __launch_bounds__(kBlock) k_halo_build(Scenes sc, Nbr nb, uint32_t limit, uint32_t* __restrict__ hidx,
                                                     uint32_t* __restrict__ hn, uint16_t* __restrict__ n16) {
    const uint32_t blk = blockIdx.x;
    const uint32_t sid = sc.blk_scene[blk];
    const StepParams P = sc.P[sid];
    const uint32_t N = P.N;
    const uint32_t tid = threadIdx.x;
    const uint32_t i = blk * kBlock + tid;
    __shared__ uint32_t key[kHaloHash];
    __shared__ uint16_t at[kHaloHash];
    __shared__ uint32_t srt[kHaloSort];
    __shared__ uint32_t scan[kBlock];
    __shared__ uint32_t cnt;
    for (uint32_t t = tid; t < kHaloHash; t += kBlock) key[t] = kNoKey;
    if (tid == 0) cnt = 0;
    __syncthreads();
    const uint32_t c = i < P.end ? nb.cff[i] : 0u;
    // insert; stop as soon as the block is known to exceed the limit. The count is checked
    // before every insert, so each thread adds at most one key after the count passes the
    // limit: the table holds at most limit + kBlock keys, never fills, and every probe ends.
    // List entries are loaded kBatch at a time so their latency overlaps.
    constexpr uint32_t kBatch = 8;
    bool over = false;
    for (uint32_t k0 = 0; k0 < c && !over; k0 += kBatch) {
        uint32_t jb[kBatch];
#pragma unroll
        for (uint32_t q = 0; q < kBatch; q++)
            jb[q] = k0 + q < c ? nb.ff[size_t(k0 + q) * N + i] : kNoKey;
#pragma unroll
        for (uint32_t q = 0; q < kBatch; q++) {
            const uint32_t j = jb[q];
            if (j == kNoKey) break;
            if (*(volatile uint32_t*)&cnt > limit) {
                over = true;
                break;
            }
            uint32_t s = halo_hash(j);
            for (uint32_t probe = 0; probe < kHaloHash; probe++) {
                const uint32_t prev = atomicCAS(&key[s], kNoKey, j);
                if (prev == kNoKey) {
                    atomicAdd(&cnt, 1u);
                    break;
                }
                if (prev == j) break;
                s = (s + 1u) & (kHaloHash - 1u);
            }
        }
    }
    __syncthreads();
    const uint32_t H = cnt;
    if (H > limit) {
        if (tid == 0) hn[blk] = kHaloCap + 1u;
        return;
    }
    // compact the occupied slots (each thread owns 16 consecutive slots) ...
    constexpr uint32_t per = kHaloHash / kBlock;
    uint32_t mine = 0;
    for (uint32_t q = 0; q < per; q++) mine += key[tid * per + q] != kNoKey;
    scan[tid] = mine;
    __syncthreads();
    for (uint32_t off = 1; off < kBlock; off <<= 1) {  // inclusive Hillis-Steele scan
        const uint32_t v = tid >= off ? scan[tid - off] : 0u;
        __syncthreads();
        scan[tid] += v;
        __syncthreads();
    }
    uint32_t o = scan[tid] - mine;
    for (uint32_t q = 0; q < per; q++) {
        const uint32_t kk = key[tid * per + q];
        if (kk != kNoKey) srt[o++] = kk;
    }
    uint32_t P2 = 1;
    while (P2 < H) P2 <<= 1;
    for (uint32_t t = H + tid; t < P2; t += kBlock) srt[t] = kNoKey;
    __syncthreads();
    // ... and sort them ascending (bitonic network over P2 <= kHaloSort keys)
    for (uint32_t size = 2; size <= P2; size <<= 1) {
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            for (uint32_t t = tid; t < P2 / 2; t += kBlock) {
                const uint32_t lo = 2u * t - (t & (stride - 1u));
                const uint32_t hi = lo + stride;
                const uint32_t a = srt[lo], b = srt[hi];
                if ((a > b) == ((lo & size) == 0u)) {
                    srt[lo] = b;
                    srt[hi] = a;
                }
            }
            __syncthreads();
        }
    }
    uint32_t* out = hidx + size_t(blk) * kHaloCap;
    for (uint32_t t = tid; t < H; t += kBlock) {
        const uint32_t j = srt[t];
        out[t] = j;
        at[halo_find(key, j)] = static_cast<uint16_t>(t);
    }
    if (tid == 0) hn[blk] = H;
    __syncthreads();
    for (uint32_t k0 = 0; k0 < c; k0 += kBatch) {
        uint32_t jb[kBatch];
#pragma unroll
        for (uint32_t q = 0; q < kBatch; q++)
            if (k0 + q < c) jb[q] = nb.ff[size_t(k0 + q) * N + i];
#pragma unroll
        for (uint32_t q = 0; q < kBatch; q++)
            if (k0 + q < c) n16[size_t(k0 + q) * N + i] = at[halo_find(key, jb[q])];
    }
}

// stats[0] = largest halo staged, stats[1] = blocks not staged, stats[2] = blocks with a halo
// or not staged, stats[3] = blocks whose halo is staged but deeper than kHaloShallow: one
// workgroup folds the per-block counts (no same-address atomics from every workgroup of the
// build).
constexpr int kHaloStatThreads = 1024;
__global__ void __launch_bounds__(kHaloStatThreads) k_halo_stats(const uint32_t* __restrict__ hn, uint32_t nblk,
                                                                uint32_t* __restrict__ stats) {
    __shared__ uint32_t wm[kHaloStatThreads / 64], ws[kHaloStatThreads / 64], wb[kHaloStatThreads / 64],
        wd[kHaloStatThreads / 64];
    uint32_t mx = 0u, sp = 0u, bl = 0u, dp = 0u;
    for (uint32_t b = threadIdx.x; b < nblk; b += blockDim.x) {
        const uint32_t h = hn[b];
        if (h > kHaloCap) sp++;
        else mx = max(mx, h);
        if (h) bl++;
        if (h > kHaloShallow && h <= kHaloCap) dp++;
    }
    for (int off = 32; off > 0; off >>= 1) {
        mx = max(mx, (uint32_t)__shfl_xor((int)mx, off, 64));
        sp += (uint32_t)__shfl_xor((int)sp, off, 64);
        bl += (uint32_t)__shfl_xor((int)bl, off, 64);
        dp += (uint32_t)__shfl_xor((int)dp, off, 64);
    }
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    if (lane == 0) {
        wm[wave] = mx;
        ws[wave] = sp;
        wb[wave] = bl;
        wd[wave] = dp;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (uint32_t w = 1; w < blockDim.x / 64; w++) {
            mx = max(mx, wm[w]);
            sp += ws[w];
            bl += wb[w];
            dp += wd[w];
        }
        stats[0] = mx;
        stats[1] = sp;
        stats[2] = bl;
        stats[3] = dp;
    }
}

}  // namespace dev

namespace launch {
void halo_build(const Scenes& sc, const Nbr& nb, uint32_t nblk, uint32_t limit, uint32_t* hidx, uint32_t* hn, uint16_t* n16,
                uint32_t* stats, hipStream_t s) {
    limit = limit == 0u || limit > kHaloCap ? kHaloCap : limit;
    if (!nblk) return;
    launch_k(dev::k_halo_build, nblk, dev::kBlock, 0, s, sc, nb, limit, hidx, hn, n16);
    launch_k(dev::k_halo_stats, 1, dev::kHaloStatThreads, 0, s, hn, nblk, stats);
}
}  // namespace launch
}  // namespace mps
