// Photon-beam LBVH build for gfx950: Morton codes, stable LSD radix sort,
// Karras hierarchy and bottom-up AABB refit — the method of the reference's
// PPB_PT.comp:317-333 (Morton), PPB_RadixSort_{LocalSort,PrefixSum,
// GlobalSort}.comp (sort), PPB_GenerateHierarchy.comp (hierarchy) and
// PPB_CalculateAABB.comp (refit), with the reference's defects corrected (each
// kernel says how).
//
// Radix sort: 4 passes of 8-bit digits.  Each pass is
//   (1) tile histograms: 4096 keys per 256-thread block, per-wave LDS
//       histograms, written tile-major (tile * 256 + digit, coalesced);
//   (2) exclusive scan of the counts: one workgroup per digit scans that
//       digit's tiles in place and writes the digit's total;
//   (3) stable scatter: each lane holds 16 of the tile's keys (one per round
//       of 256, coalesced loads); the tile's digit histogram and its scan give
//       each digit's local start; within a wave, lanes with equal digits are
//       found with 8 ballots (a wave-level multisplit), their rank is the
//       popcount of equal-digit lanes below; across the 4 waves of a round
//       and the earlier rounds, ranks are offset by LDS counters; a digit's
//       run of the tile starts at the scanned count plus the totals of the
//       smaller digits (scanned by every tile's workgroup).  The keys
//       are placed in LDS in digit order, then written out in that order, so
//       consecutive lanes write consecutive slots of each digit's run.  The
//       result is the unique stable order "by key, ties by input index".
// Hist and scatter map workgroups to tiles, and the scan to digits, so that
// data sharing a cache line is written through one XCD's L2.
// Hierarchy: one lane per internal node, Karras 2012 Fig. 4 with the
// index tie-break for equal codes; it records each node's children, leaf
// range end and parent in one 16-B record.
// Refit: one lane per leaf walks up; the second lane to reach a node unions
// the two child boxes — in LDS inside blocks of 2048 consecutive leaves, then
// in global memory through the nodes spanning blocks (two more launches).
#include "cr_device.h"
#include "cr_internal_ppb.h"

using namespace crd;

namespace {

constexpr uint32_t NONE = 0xFFFFFFFFu;
#ifndef LBVH_SORT_TILE
#define LBVH_SORT_TILE 4096
#endif
constexpr int TILE = LBVH_SORT_TILE;  // keys per tile: digit runs of 16 keys on average
constexpr uint32_t XCDS = 8;          // workgroup b runs on XCD b % 8 (round-robin dispatch)

// The tile of workgroup b: XCD x takes the x-th contiguous eighth of the
// tiles, in order, so that neighbouring tiles (whose runs of a digit are
// neighbours in the output) are written through the same L2.  The launch has
// XCDS * ceil(tiles / XCDS) workgroups; those past the last tile do nothing.
__device__ __forceinline__ uint32_t tile_of(uint32_t b, uint32_t tiles)
{
    const uint32_t per = (tiles + XCDS - 1) / XCDS;
    return (b % XCDS) * per + b / XCDS;
}
constexpr int ROUNDS = TILE / 256;

__device__ __forceinline__ uint32_t expand_bits10(uint32_t v)
{
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

__device__ __forceinline__ uint32_t morton30(f3 c, const float* b0, const float* b1)
{
    uint32_t q[3];
    const float cc[3] = {c.x, c.y, c.z};
#pragma unroll
    for (int k = 0; k < 3; k++) {
        float n = fdiv(cc[k] - b0[k], b1[k] - b0[k]);
        q[k] = (uint32_t)fminf(fmaxf(n * 1024.0f, 0.0f), 1023.0f);
    }
    return (expand_bits10(q[0]) << 2) + (expand_bits10(q[1]) << 1) + expand_bits10(q[2]);
}

__global__ void __launch_bounds__(256) morton_kernel(const cr_photon_beam* beams, uint32_t n, cr::BoxArg box, uint32_t* keys,
                                                     uint32_t* vals)
{
    uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float4* p = reinterpret_cast<const float4*>(beams + i);
    float4 a = p[0];  // startPos, mortonCode
    float4 b = p[1];  // endPos, radius
    f3 c = mk(fdiv(a.x + b.x, 2.0f), fdiv(a.y + b.y, 2.0f), fdiv(a.z + b.z, 2.0f));
    keys[i] = morton30(c, box.b0, box.b1);
    vals[i] = i;
}

// (1) per-tile digit histograms, tile-major.
__global__ void __launch_bounds__(256) hist_kernel(const uint32_t* keys, uint32_t n, uint32_t shift, uint32_t tiles,
                                                   uint32_t* counts)
{
    __shared__ uint32_t h[4][256];
    const uint32_t tile = tile_of(blockIdx.x, tiles);
    if (tile >= tiles) return;
    const uint32_t w = threadIdx.x >> 6;
    for (uint32_t i = threadIdx.x; i < 1024; i += 256) (&h[0][0])[i] = 0;
    __syncthreads();
    const uint32_t base = tile * TILE;
#pragma unroll 4
    for (int r = 0; r < ROUNDS; r++) {
        uint32_t i = base + (uint32_t)r * 256 + threadIdx.x;
        if (i < n) atomicAdd(&h[w][(keys[i] >> shift) & 0xffu], 1u);
    }
    __syncthreads();
    const uint32_t d = threadIdx.x;
    counts[(size_t)tile * 256 + d] = h[0][d] + h[1][d] + h[2][d] + h[3][d];
}

// Inclusive prefix sum over a 256-thread block (4 waves).
__device__ __forceinline__ uint32_t block_scan256(uint32_t v, uint32_t* wsum)
{
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t o = (uint32_t)__shfl_up((int)v, off, 64);
        if (lane >= (uint32_t)off) v += o;
    }
    if (lane == 63) wsum[w] = v;
    __syncthreads();
    for (uint32_t q = 0; q < w; q++) v += wsum[q];
    __syncthreads();
    return v;
}

// (2) exclusive scan of the tile-major counts in place: one block per digit
// scans the tiles' counts of that digit (a column, stride 256) and writes the
// digit's total.  The 32 digits of a 128-B line of counts are scanned on one
// XCD (block b takes digit 32 (b % 8) + b / 8), so the line is written
// through one L2.
__global__ void __launch_bounds__(256) digit_scan_kernel(uint32_t* counts, uint32_t tiles, uint32_t* totals)
{
    __shared__ uint32_t wsum[4];
    __shared__ uint32_t carry;
    const uint32_t d = (blockIdx.x % XCDS) * (256u / XCDS) + blockIdx.x / XCDS, t = threadIdx.x;
    uint32_t run = 0;
    for (uint32_t c0 = 0; c0 < tiles; c0 += 256) {
        const uint32_t i = c0 + t;
        const uint32_t x = i < tiles ? counts[(size_t)i * 256 + d] : 0u;
        const uint32_t incl = block_scan256(x, wsum);
        if (i < tiles) counts[(size_t)i * 256 + d] = run + incl - x;
        if (t == 255) carry = run + incl;
        __syncthreads();
        run = carry;
        __syncthreads();
    }
    if (t == 0) totals[d] = run;
}

// (3) stable scatter of one tile, staged in LDS in digit order.
__global__ void __launch_bounds__(256) scatter_kernel(const uint32_t* keys, const uint32_t* vals, uint32_t n, uint32_t shift,
                                                      uint32_t tiles, const uint32_t* offsets, const uint32_t* totals,
                                                      uint32_t* keys_out, uint32_t* vals_out)
{
    __shared__ uint32_t sk[TILE];       // the tile's keys and values in digit order
    __shared__ uint32_t sv[TILE];
    __shared__ uint32_t lstart[256];    // tile-local start of each digit
    __shared__ uint32_t gstart[256];    // global start of the tile's run of each digit
    __shared__ uint32_t run[256];       // next local slot of each digit
    __shared__ uint32_t wcnt[4][256];   // per-wave digit counts of the current round
    __shared__ uint32_t wsum[4];
    const uint32_t t = threadIdx.x, w = t >> 6, lane = t & 63u;
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const uint32_t tile = tile_of(blockIdx.x, tiles);
    if (tile >= tiles) return;
    const uint32_t base = tile * TILE;
    run[t] = 0;
    for (uint32_t i = t; i < 1024; i += 256) (&wcnt[0][0])[i] = 0;
    __syncthreads();
    uint32_t kk[ROUNDS], vv[ROUNDS];
#pragma unroll
    for (int r = 0; r < ROUNDS; r++) {
        const uint32_t i = base + (uint32_t)r * 256 + t;
        kk[r] = 0;
        vv[r] = 0;
        if (i < n) {
            kk[r] = keys[i];
            vv[r] = vals[i];
            atomicAdd(&run[(kk[r] >> shift) & 0xffu], 1u);
        }
    }
    __syncthreads();
    const uint32_t c = run[t];
    const uint32_t incl = block_scan256(c, wsum);
    const uint32_t tot = totals[t];
    const uint32_t below = block_scan256(tot, wsum) - tot;  // keys of the smaller digits
    lstart[t] = incl - c;
    run[t] = incl - c;
    gstart[t] = below + offsets[(size_t)tile * 256 + t];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < ROUNDS; r++) {
        const uint32_t i = base + (uint32_t)r * 256 + t;
        const bool valid = i < n;
        const uint32_t d = valid ? (kk[r] >> shift) & 0xffu : 0u;
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const bool bit = (d >> b) & 1u;
            const unsigned long long bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        const uint32_t rank = (uint32_t)__popcll(peers & lt);
        if (valid && rank == 0) wcnt[w][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t pos = run[d] + rank;
            for (uint32_t q = 0; q < w; q++) pos += wcnt[q][d];
            sk[pos] = kk[r];
            sv[pos] = vv[r];
        }
        __syncthreads();
        run[t] += wcnt[0][t] + wcnt[1][t] + wcnt[2][t] + wcnt[3][t];
        wcnt[0][t] = 0; wcnt[1][t] = 0; wcnt[2][t] = 0; wcnt[3][t] = 0;
        __syncthreads();
    }
    const uint32_t cnt = min((uint32_t)TILE, n - base);
    for (uint32_t j = t; j < cnt; j += 256) {
        const uint32_t k = sk[j];
        const uint32_t d = (k >> shift) & 0xffu;
        const uint32_t pos = gstart[d] + (j - lstart[d]);
        keys_out[pos] = k;
        vals_out[pos] = sv[j];
    }
}

// delta(i, j) of Karras 2012 with the index tie-break (exact integer clz; equal codes split by index).
__device__ __forceinline__ int delta(const uint32_t* k, int32_t n, int32_t i, int32_t j)
{
    if (j < 0 || j > n - 1) return -1;
    const uint32_t a = k[i], b = k[j];
    if (a == b) return 32 + __clz((int)((uint32_t)i ^ (uint32_t)j));
    return __clz((int)(a ^ b));
}

// meta[i] = (left, right, end, parent) of internal node i: children as node
// indices (leaf l is node n - 1 + l), the other end of its leaf range
// [min(i, end), max(i, end)], and its parent (NONE at the root), which the
// parent's lane writes.  leafParent[l]: the parent of leaf l.
__global__ void __launch_bounds__(256) hierarchy_kernel(const uint32_t* k, int32_t n, uint4* meta, uint32_t* leafParent)
{
    const int32_t i = (int32_t)(blockIdx.x * 256 + threadIdx.x);
    if (i >= n - 1) return;
    const int d = (delta(k, n, i, i + 1) - delta(k, n, i, i - 1)) > 0 ? 1 : -1;
    const int dmin = delta(k, n, i, i - d);
    int32_t lmax = 2;
    while (delta(k, n, i, i + lmax * d) > dmin) lmax *= 2;
    int32_t l = 0;
    for (int32_t t = lmax / 2; t >= 1; t /= 2)
        if (delta(k, n, i, i + (l + t) * d) > dmin) l += t;
    const int32_t j = i + l * d;
    const int32_t first = min(i, j), last = max(i, j);
    // findSplit on (first, last)
    const int dnode = delta(k, n, first, last);
    int32_t s = first, step = last - first;
    do {
        step = (step + 1) >> 1;
        const int32_t ns = s + step;
        if (ns < last && delta(k, n, first, ns) > dnode) s = ns;
    } while (step > 1);
    const uint32_t inner = (uint32_t)(n - 1);
    const uint32_t lc = (first == s) ? inner + (uint32_t)s : (uint32_t)s;
    const uint32_t rc = (last == s + 1) ? inner + (uint32_t)(s + 1) : (uint32_t)(s + 1);
    uint32_t* mi = reinterpret_cast<uint32_t*>(meta + i);
    mi[0] = lc;
    mi[1] = rc;
    mi[2] = (uint32_t)j;  // node i covers leaves [min(i, j), max(i, j)]
    if (lc < inner) reinterpret_cast<uint32_t*>(meta + lc)[3] = (uint32_t)i;
    else leafParent[lc - inner] = (uint32_t)i;
    if (rc < inner) reinterpret_cast<uint32_t*>(meta + rc)[3] = (uint32_t)i;
    else leafParent[rc - inner] = (uint32_t)i;
    if (i == 0) mi[3] = NONE;
}

// Reorder the beams into sorted order (the reference's sort moves whole
// PhotonBeam records; here the 36 bytes the gather reads: the segment's ends,
// radius and data index, and its offset along its parent beam) and write each
// leaf's box (calculateBeamBounds, PPB_CalculateAABB.comp:64-80; a
// zero-length beam gets the ball of its radius).
__global__ void __launch_bounds__(256) leaf_kernel(const cr_photon_beam* beams, const uint32_t* perm, uint32_t n,
                                                   cr::SortedBeam* sorted, float* sortedOff, float* lo, float* hi)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const cr_photon_beam b = beams[perm[i]];
    cr::SortedBeam sb;
    sb.startPos[0] = b.startPos[0]; sb.startPos[1] = b.startPos[1]; sb.startPos[2] = b.startPos[2];
    sb.radius = b.radius;
    sb.endPos[0] = b.endPos[0]; sb.endPos[1] = b.endPos[1]; sb.endPos[2] = b.endPos[2];
    sb.dataIdx = b.dataIdx;
    sorted[i] = sb;
    sortedOff[i] = __uint_as_float(b._padding_beam[0]);
    f3 pa = ld3(b.startPos), pb = ld3(b.endPos);
    f3 a = sub(pb, pa);
    float dd = dot(a, a);
    float ra = b.radius;
    f3 e;
    if (dd > 0.0f) {
        e = mk(ra * fsqrt(1.0f - fdiv(a.x * a.x, dd)), ra * fsqrt(1.0f - fdiv(a.y * a.y, dd)),
               ra * fsqrt(1.0f - fdiv(a.z * a.z, dd)));
    } else {
        e = mk(ra, ra, ra);
    }
    const size_t L = (size_t)(n - 1) + i;
    lo[3 * L + 0] = fminf(pa.x - e.x, pb.x - e.x);
    lo[3 * L + 1] = fminf(pa.y - e.y, pb.y - e.y);
    lo[3 * L + 2] = fminf(pa.z - e.z, pb.z - e.z);
    hi[3 * L + 0] = fmaxf(pa.x + e.x, pb.x + e.x);
    hi[3 * L + 1] = fmaxf(pa.y + e.y, pb.y + e.y);
    hi[3 * L + 2] = fmaxf(pa.z + e.z, pb.z + e.z);
}

// Bottom-up union (PPB_CalculateAABB.comp:100-120, corrected): the first
// lane to arrive at a node stops; the second unions both children.  Internal
// node i covers leaves [min(i, end), max(i, end)] (meta[i].z), and a node
// whose leaves all lie in a range of leaves has its index in that range and
// both subtrees inside it.  Three launches, so no lane ever waits for one on
// another XCD and most of the climbing happens in LDS:
//  * refit_block_kernel: a workgroup takes REFIT_BLOCK consecutive leaves and
//    holds their boxes, the boxes and records of the internal nodes with
//    indices in the block, and their arrival counters in LDS; it climbs from
//    every leaf through the nodes contained in the block, then writes those
//    nodes' boxes out (coalesced).  A lane that reaches a node spanning
//    several blocks records it in the block's exit list (in LDS, written out
//    after the climb: no global atomics) and stops.
//  * refit_mid_kernel: a workgroup per REFIT_SUPER consecutive leaves climbs,
//    in global memory with workgroup-scope counters, from the exits of its
//    blocks through the nodes contained in its range; nodes not contained in
//    it (their index may lie outside it) go to the second exit list.
//  * refit_top_kernel: one workgroup climbs from the second list to the root.
// Every node's two arrivals happen in the same launch (the one whose range
// first contains the node's leaves), and each launch's counters are touched
// by one workgroup only.
constexpr uint32_t REFIT_BLOCK = 2048;                // leaves per block: 139,264 B of LDS
constexpr uint32_t REFIT_SUPER = 64 * REFIT_BLOCK;    // leaves per mid-level range
constexpr int REFIT_THREADS = 1024;

// exits: per block REFIT_BLOCK entries (block k's at k * REFIT_BLOCK), then
// the blocks' entry counts.
__global__ void __launch_bounds__(REFIT_THREADS) refit_block_kernel(uint32_t n, const uint4* meta, const uint32_t* leafParent,
                                                                    float* lo, float* hi, uint32_t* exits, uint32_t* exitCounts)
{
    // slots: internal node i -> i - first; leaf l -> REFIT_BLOCK + l - first
    __shared__ float blo[2 * REFIT_BLOCK][3];
    __shared__ float bhi[2 * REFIT_BLOCK][3];
    __shared__ uint4 smeta[REFIT_BLOCK];
    __shared__ uint32_t arrived[REFIT_BLOCK];
    __shared__ uint32_t sexit[REFIT_BLOCK];  // each leaf's climb exits at most once
    __shared__ uint32_t nexit;
    const uint32_t inner = n - 1;
    const uint32_t first = blockIdx.x * REFIT_BLOCK;
    const uint32_t end = min(first + REFIT_BLOCK, n);          // leaves [first, end)
    const uint32_t iend = min(first + REFIT_BLOCK, inner);     // internal nodes [first, iend)
    if (threadIdx.x == 0) nexit = 0u;
    for (uint32_t t = threadIdx.x; t < REFIT_BLOCK; t += REFIT_THREADS) {
        arrived[t] = 0u;
        if (first + t < iend) smeta[t] = meta[first + t];
        if (first + t < end) {
            const size_t L = (size_t)inner + first + t;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                blo[REFIT_BLOCK + t][k] = lo[3 * L + k];
                bhi[REFIT_BLOCK + t][k] = hi[3 * L + k];
            }
        }
    }
    __syncthreads();
    for (uint32_t leaf = first + threadIdx.x; leaf < end; leaf += REFIT_THREADS) {
        uint32_t cur = leafParent[leaf];
        // a Karras tree over 30-bit codes with index tie-breaks is at most 64
        // levels deep; the bound only protects against a malformed tree
        for (int level = 0; level < 128; level++) {
            const bool inBlock = cur >= first && cur < iend;
            const uint4 mc = inBlock ? smeta[cur - first] : make_uint4(0u, 0u, 0u, 0u);
            if (!inBlock || !(min(cur, mc.z) >= first && max(cur, mc.z) < end)) {
                sexit[atomicAdd(&nexit, 1u)] = cur;
                break;
            }
            const uint32_t old =
                __hip_atomic_fetch_add(&arrived[cur - first], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (old == 0u) break;
            const uint32_t sa = mc.x < inner ? mc.x - first : REFIT_BLOCK + (mc.x - inner - first);
            const uint32_t sb = mc.y < inner ? mc.y - first : REFIT_BLOCK + (mc.y - inner - first);
            const uint32_t sc = cur - first;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                blo[sc][k] = fminf(blo[sa][k], blo[sb][k]);
                bhi[sc][k] = fmaxf(bhi[sa][k], bhi[sb][k]);
            }
            if (!(mc.w < inner)) break;  // the root
            cur = mc.w;
        }
    }
    __syncthreads();
    // the boxes of the internal nodes contained in the block (both children arrived)
    for (uint32_t t = threadIdx.x; first + t < iend; t += REFIT_THREADS) {
        if (arrived[t] == 2u) {
#pragma unroll
            for (int k = 0; k < 3; k++) {
                lo[3 * (size_t)(first + t) + k] = blo[t][k];
                hi[3 * (size_t)(first + t) + k] = bhi[t][k];
            }
        }
    }
    const uint32_t ne = nexit;
    for (uint32_t t = threadIdx.x; t < ne; t += REFIT_THREADS) exits[(size_t)blockIdx.x * REFIT_BLOCK + t] = sexit[t];
    if (threadIdx.x == 0) exitCounts[blockIdx.x] = ne;
}

// One node's arrival in global memory (workgroup-scope counter): false for
// the first arrival; the second forms the node's box from its children's.
__device__ __forceinline__ bool arrive_global(uint32_t cur, const uint4& mc, float* lo, float* hi, uint32_t* arrivals)
{
    const uint32_t old = __hip_atomic_fetch_add(arrivals + cur, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (old == 0u) return false;
    const uint32_t a = mc.x, b = mc.y;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        lo[3 * (size_t)cur + k] = fminf(lo[3 * (size_t)a + k], lo[3 * (size_t)b + k]);
        hi[3 * (size_t)cur + k] = fmaxf(hi[3 * (size_t)a + k], hi[3 * (size_t)b + k]);
    }
    return true;
}

// Sixteen lanes per block of the range take that block's exits.
constexpr uint32_t REFIT_SUPER_BLOCKS = REFIT_SUPER / REFIT_BLOCK;
static_assert(REFIT_SUPER_BLOCKS * 16 == REFIT_THREADS, "sixteen lanes per block");

__global__ void __launch_bounds__(REFIT_THREADS) refit_mid_kernel(uint32_t n, const uint4* meta, float* lo, float* hi,
                                                                  const uint32_t* exits, const uint32_t* exitCounts,
                                                                  uint32_t* arrivals, uint32_t* exits2, uint32_t* exitCount2)
{
    const uint32_t inner = n - 1;
    const uint32_t first = blockIdx.x * REFIT_SUPER;
    const uint32_t last = min(first + REFIT_SUPER, n) - 1u;  // leaves [first, last]
    const uint32_t block = blockIdx.x * REFIT_SUPER_BLOCKS + threadIdx.x / 16u;
    const uint32_t nBlocks = (n + REFIT_BLOCK - 1) / REFIT_BLOCK;
    const uint32_t count = block < nBlocks ? exitCounts[block] : 0u;
    for (uint32_t x = threadIdx.x % 16u; x < count; x += 16u) {
        uint32_t cur = exits[(size_t)block * REFIT_BLOCK + x];
        for (int level = 0; level < 128; level++) {
            const uint4 mc = meta[cur];
            if (!(min(cur, mc.z) >= first && max(cur, mc.z) <= last)) {
                exits2[atomicAdd(exitCount2, 1u)] = cur;
                break;
            }
            if (!arrive_global(cur, mc, lo, hi, arrivals)) break;
            if (!(mc.w < inner)) break;
            cur = mc.w;
        }
    }
}

constexpr int REFIT_TOP_THREADS = 1024;

__global__ void __launch_bounds__(REFIT_TOP_THREADS) refit_top_kernel(uint32_t n, const uint4* meta, float* lo, float* hi,
                                                                      const uint32_t* exits, const uint32_t* exitCount,
                                                                      uint32_t* arrivals)
{
    const uint32_t inner = n - 1;
    const uint32_t count = *exitCount;
    for (uint32_t x = threadIdx.x; x < count; x += REFIT_TOP_THREADS) {
        uint32_t cur = exits[x];
        for (int level = 0; cur < inner && level < 128; level++) {
            const uint4 mc = meta[cur];
            if (!arrive_global(cur, mc, lo, hi, arrivals)) break;
            cur = mc.w;
        }
    }
}

// Traversal layout: one 64-B record per internal node holding both child
// boxes and child indices, so one traversal step reads one cache line.
__global__ void __launch_bounds__(256) pack_kernel(uint32_t n, const uint4* meta, const float* lo, const float* hi,
                                                   cr::TravNode* out)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i + 1 >= n) return;
    const uint4 mi = meta[i];
    const uint32_t a = mi.x, b = mi.y;
    cr::TravNode t;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const float la = lo[3 * (size_t)a + k], ha = hi[3 * (size_t)a + k];
        const float lb = lo[3 * (size_t)b + k], hb = hi[3 * (size_t)b + k];
        // (lo + hi) / 2 and (hi - lo) * 0.5: halving is exact, so * 0.5 is the same value
        t.cen[0][k] = (la + ha) * 0.5f;
        t.rad[0][k] = (ha - la) * 0.5f;
        t.cen[1][k] = (lb + hb) * 0.5f;
        t.rad[1][k] = (hb - lb) * 0.5f;
    }
    t.child[0] = a;
    t.child[1] = b;
    t.leafMask = (a >= n - 1 ? 1u : 0u) | (b >= n - 1 ? 2u : 0u);
    t.pad = 0;
    out[i] = t;
}

// Reference TreeNode layout (UniformBuffers.h:179-188) for download.
__global__ void __launch_bounds__(256) treenode_kernel(uint32_t n, const uint4* meta, const uint32_t* leafParent,
                                                       const float* lo, const float* hi, cr_tree_node* out)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    const uint32_t total = 2 * n - 1;
    if (i >= total) return;
    const uint32_t inner = n - 1;
    cr_tree_node t;
    const bool leaf = i >= inner;
    const uint4 mi = leaf ? make_uint4(0u, 0u, 0u, 0u) : meta[i];
    t.parent = leaf ? leafParent[i - inner] : mi.w;
    t.left = leaf ? i - inner : mi.x;
    t.right = leaf ? NONE : mi.y;
    t.isLeaf = leaf ? 1u : 0u;
    for (int k = 0; k < 3; k++) {
        t.bounds[0][k] = lo[3 * (size_t)i + k];
        t.bounds[1][k] = hi[3 * (size_t)i + k];
    }
    t.bounds[0][3] = 0.0f;
    t.bounds[1][3] = 0.0f;
    t.processed = leaf ? 0u : 2u;
    t._padding_node[0] = t._padding_node[1] = t._padding_node[2] = 0u;
    out[i] = t;
}

inline unsigned blocks_for(uint64_t n, unsigned per) { return (unsigned)((n + per - 1) / per); }

} // namespace

namespace cr {

hipError_t launch_morton(const cr_photon_beam* beams, uint32_t n, const BoxArg& box, uint32_t* keys, uint32_t* vals,
                         hipStream_t s)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(morton_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, beams, n, box, keys, vals);
    return hipGetLastError();
}

uint32_t sort_tiles(uint32_t n) { return (n + TILE - 1) / TILE; }
uint32_t sort_counts_words(uint32_t n) { return sort_tiles(n) * 256u + 256u; }

hipError_t launch_radix_sort(uint32_t* keys, uint32_t* vals, uint32_t* keys_tmp, uint32_t* vals_tmp, uint32_t* counts,
                             uint32_t n, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    const uint32_t tiles = sort_tiles(n);
    uint32_t* ki = keys;
    uint32_t* vi = vals;
    uint32_t* ko = keys_tmp;
    uint32_t* vo = vals_tmp;
    // the digit totals follow the histograms
    uint32_t* totals = counts + (size_t)tiles * 256u;
    const uint32_t grid = (tiles + XCDS - 1) / XCDS * XCDS;  // see tile_of
    for (uint32_t pass = 0; pass < 4; pass++) {
        const uint32_t shift = 8u * pass;
        hipLaunchKernelGGL(hist_kernel, dim3(grid), dim3(256), 0, s, ki, n, shift, tiles, counts);
        hipLaunchKernelGGL(digit_scan_kernel, dim3(256), dim3(256), 0, s, counts, tiles, totals);
        hipLaunchKernelGGL(scatter_kernel, dim3(grid), dim3(256), 0, s, ki, vi, n, shift, tiles, counts, totals, ko, vo);
        uint32_t* t = ki; ki = ko; ko = t;
        t = vi; vi = vo; vo = t;
    }
    // 4 passes: the result is back in keys/vals
    return hipGetLastError();
}

hipError_t launch_hierarchy(const uint32_t* keys, uint32_t n, uint4* meta, uint32_t* leafParent, hipStream_t s)
{
    if (n < 2) return hipSuccess;
    hipLaunchKernelGGL(hierarchy_kernel, dim3(blocks_for(n - 1, 256)), dim3(256), 0, s, keys, (int32_t)n, meta, leafParent);
    return hipGetLastError();
}

hipError_t launch_leaves(const cr_photon_beam* beams, const uint32_t* perm, uint32_t n, SortedBeam* sorted, float* sortedOff,
                         float* lo, float* hi, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(leaf_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, beams, perm, n, sorted, sortedOff, lo, hi);
    return hipGetLastError();
}

// exits: the blocks' lists (REFIT_BLOCK entries each) and counts, then the
// second list (at most two entries per internal node) and its count, last.
uint64_t refit_exit_words(uint32_t n)
{
    const uint64_t nb = (n + REFIT_BLOCK - 1) / REFIT_BLOCK;
    return nb * REFIT_BLOCK + nb + 2ull * n + 1ull;
}

hipError_t launch_refit(uint32_t n, const uint4* meta, const uint32_t* leafParent, float* lo, float* hi, uint32_t* arrivals,
                        uint32_t* exits, hipStream_t s)
{
    if (n < 2) return hipSuccess;
    const uint64_t nb = (n + REFIT_BLOCK - 1) / REFIT_BLOCK;
    uint32_t* count1 = exits + nb * REFIT_BLOCK;
    uint32_t* list2 = count1 + nb;
    uint32_t* count2 = exits + refit_exit_words(n) - 1;
    hipLaunchKernelGGL(refit_block_kernel, dim3((unsigned)nb), dim3(REFIT_THREADS), 0, s, n, meta, leafParent, lo, hi, exits,
                       count1);
    hipLaunchKernelGGL(refit_mid_kernel, dim3(blocks_for(n, REFIT_SUPER)), dim3(REFIT_THREADS), 0, s, n, meta, lo, hi, exits,
                       count1, arrivals, list2, count2);
    hipLaunchKernelGGL(refit_top_kernel, dim3(1), dim3(REFIT_TOP_THREADS), 0, s, n, meta, lo, hi, list2, count2, arrivals);
    return hipGetLastError();
}

hipError_t launch_pack(uint32_t n, const uint4* meta, const float* lo, const float* hi, TravNode* out, hipStream_t s)
{
    if (n < 2) return hipSuccess;
    hipLaunchKernelGGL(pack_kernel, dim3(blocks_for(n - 1, 256)), dim3(256), 0, s, n, meta, lo, hi, out);
    return hipGetLastError();
}

hipError_t launch_treenodes(uint32_t n, const uint4* meta, const uint32_t* leafParent, const float* lo, const float* hi,
                            cr_tree_node* out, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(treenode_kernel, dim3(blocks_for(2ull * n - 1, 256)), dim3(256), 0, s, n, meta, leafParent, lo, hi, out);
    return hipGetLastError();
}

} // namespace cr
