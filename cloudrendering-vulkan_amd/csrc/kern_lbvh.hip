// Photon-beam LBVH build for gfx950: Morton codes, stable LSD radix sort,
// Karras hierarchy and bottom-up AABB refit — the method of the reference's
// PPB_PT.comp:317-333 (Morton), PPB_RadixSort_{LocalSort,PrefixSum,
// GlobalSort}.comp (sort), PPB_GenerateHierarchy.comp (hierarchy) and
// PPB_CalculateAABB.comp (refit), with the reference's defects corrected (each
// kernel says how).
//
// Radix sort: 4 passes of 8-bit digits.  Each pass is
//   (1) tile histograms: 4096 keys per 256-thread block, per-wave LDS
//       histograms, written digit-major (digit * tiles + tile), and each
//       digit's total (global atomics);
//   (2) exclusive scan of the counts: one workgroup per digit scans that
//       digit's tiles, offset by the totals of the smaller digits;
//   (3) stable scatter: the tile is consumed in rounds of 256 keys (one key
//       per lane, coalesced); within a wave, lanes with equal digits are found
//       with 8 ballots (a wave-level multisplit), their rank is the popcount
//       of equal-digit lanes below; across the 4 waves of the round and the
//       previous rounds of the tile, ranks are offset by LDS counters.  The
//       result is the unique stable order "by key, ties by input index".
// Hierarchy: one lane per internal node, Karras 2012 Fig. 4 with the
// index tie-break for equal codes; it also records each node's leaf range.
// Refit: one lane per leaf walks up; the second lane to reach a node unions
// the two child boxes — first inside workgroups of consecutive leaves (LDS
// counters), then, in a second launch, through the nodes spanning them.
#include "cr_device.h"
#include "cr_internal_ppb.h"

using namespace crd;

namespace {

constexpr uint32_t NONE = 0xFFFFFFFFu;
constexpr int TILE = 4096;
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

// (1) per-tile digit histograms, digit-major, and each digit's total.
__global__ void __launch_bounds__(256) hist_kernel(const uint32_t* keys, uint32_t n, uint32_t shift, uint32_t tiles,
                                                   uint32_t* counts, uint32_t* totals)
{
    __shared__ uint32_t h[4][256];
    const uint32_t w = threadIdx.x >> 6;
    for (uint32_t i = threadIdx.x; i < 1024; i += 256) (&h[0][0])[i] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * TILE;
#pragma unroll 4
    for (int r = 0; r < ROUNDS; r++) {
        uint32_t i = base + (uint32_t)r * 256 + threadIdx.x;
        if (i < n) atomicAdd(&h[w][(keys[i] >> shift) & 0xffu], 1u);
    }
    __syncthreads();
    const uint32_t d = threadIdx.x;
    const uint32_t c = h[0][d] + h[1][d] + h[2][d] + h[3][d];
    counts[(size_t)d * tiles + blockIdx.x] = c;
    if (c) atomicAdd(totals + d, c);
}

__global__ void __launch_bounds__(256) zero_totals_kernel(uint32_t* totals) { totals[blockIdx.x * 256 + threadIdx.x] = 0u; }

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

// (2) exclusive scan of the digit-major counts in place: block d scans the
// tiles' counts of digit d, offset by the totals of the digits below d.
__global__ void __launch_bounds__(256) digit_scan_kernel(uint32_t* counts, uint32_t tiles, const uint32_t* totals)
{
    __shared__ uint32_t wsum[4];
    __shared__ uint32_t carry;
    const uint32_t d = blockIdx.x, t = threadIdx.x;
    const uint32_t below = block_scan256(t < d ? totals[t] : 0u, wsum);
    if (t == 255) carry = below;  // sum of totals[0 .. d-1]
    __syncthreads();
    uint32_t run = carry;
    uint32_t* row = counts + (size_t)d * tiles;
    for (uint32_t c0 = 0; c0 < tiles; c0 += 256) {
        const uint32_t i = c0 + t;
        const uint32_t x = i < tiles ? row[i] : 0u;
        const uint32_t incl = block_scan256(x, wsum);
        if (i < tiles) row[i] = run + incl - x;
        if (t == 255) carry = run + incl;
        __syncthreads();
        run = carry;
        __syncthreads();
    }
}

// (3) stable scatter of one tile.
__global__ void __launch_bounds__(256) scatter_kernel(const uint32_t* keys, const uint32_t* vals, uint32_t n, uint32_t shift,
                                                      uint32_t tiles, const uint32_t* offsets, uint32_t* keys_out,
                                                      uint32_t* vals_out)
{
    __shared__ uint32_t run[256];      // keys of each digit already placed by earlier rounds
    __shared__ uint32_t wcnt[4][256];  // per-wave digit counts of the current round
    const uint32_t t = threadIdx.x, w = t >> 6, lane = t & 63u;
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    run[t] = offsets[(size_t)t * tiles + blockIdx.x];
    for (uint32_t i = t; i < 1024; i += 256) (&wcnt[0][0])[i] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * TILE;
    for (int r = 0; r < ROUNDS; r++) {
        const uint32_t i = base + (uint32_t)r * 256 + t;
        const bool valid = i < n;
        uint32_t k = 0, v = 0, d = 0;
        if (valid) {
            k = keys[i];
            v = vals[i];
            d = (k >> shift) & 0xffu;
        }
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
            keys_out[pos] = k;
            vals_out[pos] = v;
        }
        __syncthreads();
        run[t] += wcnt[0][t] + wcnt[1][t] + wcnt[2][t] + wcnt[3][t];
        wcnt[0][t] = 0; wcnt[1][t] = 0; wcnt[2][t] = 0; wcnt[3][t] = 0;
        __syncthreads();
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

__global__ void __launch_bounds__(256) hierarchy_kernel(const uint32_t* k, int32_t n, uint32_t* parent, uint32_t* left,
                                                        uint32_t* right, uint32_t* ends)
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
    left[i] = lc;
    right[i] = rc;
    ends[i] = (uint32_t)j;  // node i covers leaves [min(i, j), max(i, j)]
    parent[lc] = (uint32_t)i;
    parent[rc] = (uint32_t)i;
    if (i == 0) parent[0] = NONE;
}

// Reorder the beams into sorted order (the reference's sort moves whole
// PhotonBeam records) and write each leaf's box (calculateBeamBounds,
// PPB_CalculateAABB.comp:64-80; a zero-length beam gets the ball of its radius).
__global__ void __launch_bounds__(256) leaf_kernel(const cr_photon_beam* beams, const uint32_t* perm, const uint32_t* keys,
                                                   uint32_t n, cr_photon_beam* sorted, float* lo, float* hi)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    cr_photon_beam b = beams[perm[i]];
    b.mortonCode = keys[i];
    sorted[i] = b;
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
// lane to arrive at a node stops; the second unions both children.  Two
// launches, so no lane ever waits for one on another XCD:
//  * refit_local_kernel: a workgroup takes REFIT_LEAVES consecutive leaves.
//    Internal node i covers leaves [min(i, j), max(i, j)] (j = ends[i]); a
//    node whose leaves all belong to the workgroup has both subtrees inside
//    it, so its arrival counter lives in LDS and a workgroup-scope acq_rel
//    RMW orders the box stores and loads of the workgroup's lanes.  A lane
//    that completes a child of a node spanning several workgroups appends
//    that node to `exits` and stops.
//  * refit_top_kernel: one workgroup climbs from every exit through the
//    spanning nodes (the ancestors of the workgroup boundaries, a few per
//    workgroup of the first launch), with workgroup-scope RMWs on global
//    counters; the boxes of their children are the first launch's output.
constexpr uint32_t REFIT_LEAVES = 16384;  // 64 KiB of LDS counters
constexpr int REFIT_THREADS = 1024;

__global__ void __launch_bounds__(REFIT_THREADS) refit_local_kernel(uint32_t n, const uint32_t* parent, const uint32_t* left,
                                                                    const uint32_t* right, const uint32_t* ends, float* lo,
                                                                    float* hi, uint32_t* exits, uint32_t* exitCount)
{
    __shared__ uint32_t arrived[REFIT_LEAVES];
    const uint32_t first = blockIdx.x * REFIT_LEAVES;
    for (uint32_t t = threadIdx.x; t < REFIT_LEAVES; t += REFIT_THREADS) arrived[t] = 0u;
    __syncthreads();
    const uint32_t inner = n - 1;
    for (uint32_t leaf = first + threadIdx.x; leaf < first + REFIT_LEAVES && leaf < n; leaf += REFIT_THREADS) {
        uint32_t cur = parent[inner + leaf];
        // A Karras tree over 30-bit codes with index tie-breaks is at most 64
        // levels deep; the bound only protects against a malformed tree.
        for (int level = 0; cur != NONE && cur < inner && level < 128; level++) {
            const uint32_t e = ends[cur];
            if (!(min(cur, e) >= first && max(cur, e) < first + REFIT_LEAVES)) {
                exits[atomicAdd(exitCount, 1u)] = cur;
                break;
            }
            const uint32_t old =
                __hip_atomic_fetch_add(&arrived[cur - first], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (old == 0u) break;
            const uint32_t a = left[cur], b = right[cur];
#pragma unroll
            for (int k = 0; k < 3; k++) {
                lo[3 * (size_t)cur + k] = fminf(lo[3 * (size_t)a + k], lo[3 * (size_t)b + k]);
                hi[3 * (size_t)cur + k] = fmaxf(hi[3 * (size_t)a + k], hi[3 * (size_t)b + k]);
            }
            cur = parent[cur];
        }
    }
}

constexpr int REFIT_TOP_THREADS = 1024;

__global__ void __launch_bounds__(REFIT_TOP_THREADS) refit_top_kernel(uint32_t n, const uint32_t* parent,
                                                                      const uint32_t* left, const uint32_t* right,
                                                                      float* lo, float* hi, const uint32_t* exits,
                                                                      const uint32_t* exitCount, uint32_t* arrivals)
{
    const uint32_t inner = n - 1;
    const uint32_t count = *exitCount;
    for (uint32_t x = threadIdx.x; x < count; x += REFIT_TOP_THREADS) {
        uint32_t cur = exits[x];
        for (int level = 0; cur != NONE && cur < inner && level < 128; level++) {
            const uint32_t old =
                __hip_atomic_fetch_add(arrivals + cur, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (old == 0u) break;
            const uint32_t a = left[cur], b = right[cur];
#pragma unroll
            for (int k = 0; k < 3; k++) {
                lo[3 * (size_t)cur + k] = fminf(lo[3 * (size_t)a + k], lo[3 * (size_t)b + k]);
                hi[3 * (size_t)cur + k] = fmaxf(hi[3 * (size_t)a + k], hi[3 * (size_t)b + k]);
            }
            cur = parent[cur];
        }
    }
}

// Traversal layout: one 64-B record per internal node holding both child
// boxes and child indices, so one traversal step reads one cache line.
__global__ void __launch_bounds__(256) pack_kernel(uint32_t n, const uint32_t* left, const uint32_t* right, const float* lo,
                                                   const float* hi, cr::TravNode* out)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i + 1 >= n) return;
    const uint32_t a = left[i], b = right[i];
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
__global__ void __launch_bounds__(256) treenode_kernel(uint32_t n, const uint32_t* parent, const uint32_t* left,
                                                       const uint32_t* right, const float* lo, const float* hi,
                                                       cr_tree_node* out)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    const uint32_t total = 2 * n - 1;
    if (i >= total) return;
    const uint32_t inner = n - 1;
    cr_tree_node t;
    t.parent = parent[i];
    const bool leaf = i >= inner;
    t.left = leaf ? i - inner : left[i];
    t.right = leaf ? NONE : right[i];
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
uint32_t sort_counts_words(uint32_t n) { return sort_tiles(n) * 256u + 4u * 256u; }

hipError_t launch_radix_sort(uint32_t* keys, uint32_t* vals, uint32_t* keys_tmp, uint32_t* vals_tmp, uint32_t* counts,
                             uint32_t n, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    const uint32_t tiles = sort_tiles(n);
    uint32_t* ki = keys;
    uint32_t* vi = vals;
    uint32_t* ko = keys_tmp;
    uint32_t* vo = vals_tmp;
    // each pass's digit totals follow the histograms
    uint32_t* totals = counts + (size_t)tiles * 256u;
    hipLaunchKernelGGL(zero_totals_kernel, dim3(4), dim3(256), 0, s, totals);
    for (uint32_t pass = 0; pass < 4; pass++) {
        const uint32_t shift = 8u * pass;
        uint32_t* tot = totals + 256u * pass;
        hipLaunchKernelGGL(hist_kernel, dim3(tiles), dim3(256), 0, s, ki, n, shift, tiles, counts, tot);
        hipLaunchKernelGGL(digit_scan_kernel, dim3(256), dim3(256), 0, s, counts, tiles, tot);
        hipLaunchKernelGGL(scatter_kernel, dim3(tiles), dim3(256), 0, s, ki, vi, n, shift, tiles, counts, ko, vo);
        uint32_t* t = ki; ki = ko; ko = t;
        t = vi; vi = vo; vo = t;
    }
    // 4 passes: the result is back in keys/vals
    return hipGetLastError();
}

hipError_t launch_hierarchy(const uint32_t* keys, uint32_t n, uint32_t* parent, uint32_t* left, uint32_t* right,
                            uint32_t* ends, hipStream_t s)
{
    if (n < 2) return hipSuccess;
    hipLaunchKernelGGL(hierarchy_kernel, dim3(blocks_for(n - 1, 256)), dim3(256), 0, s, keys, (int32_t)n, parent, left, right,
                       ends);
    return hipGetLastError();
}

hipError_t launch_leaves(const cr_photon_beam* beams, const uint32_t* perm, const uint32_t* keys, uint32_t n,
                         cr_photon_beam* sorted, float* lo, float* hi, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(leaf_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, beams, perm, keys, n, sorted, lo, hi);
    return hipGetLastError();
}

hipError_t launch_refit(uint32_t n, const uint32_t* parent, const uint32_t* left, const uint32_t* right, const uint32_t* ends,
                        float* lo, float* hi, uint32_t* arrivals, uint32_t* exits, uint32_t* exitCount, hipStream_t s)
{
    if (n < 2) return hipSuccess;
    hipLaunchKernelGGL(refit_local_kernel, dim3(blocks_for(n, REFIT_LEAVES)), dim3(REFIT_THREADS), 0, s, n, parent, left,
                       right, ends, lo, hi, exits, exitCount);
    hipLaunchKernelGGL(refit_top_kernel, dim3(1), dim3(REFIT_TOP_THREADS), 0, s, n, parent, left, right, lo, hi, exits,
                       exitCount, arrivals);
    return hipGetLastError();
}

hipError_t launch_pack(uint32_t n, const uint32_t* left, const uint32_t* right, const float* lo, const float* hi,
                       TravNode* out, hipStream_t s)
{
    if (n < 2) return hipSuccess;
    hipLaunchKernelGGL(pack_kernel, dim3(blocks_for(n - 1, 256)), dim3(256), 0, s, n, left, right, lo, hi, out);
    return hipGetLastError();
}

hipError_t launch_treenodes(uint32_t n, const uint32_t* parent, const uint32_t* left, const uint32_t* right, const float* lo,
                            const float* hi, cr_tree_node* out, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(treenode_kernel, dim3(blocks_for(2ull * n - 1, 256)), dim3(256), 0, s, n, parent, left, right, lo, hi, out);
    return hipGetLastError();
}

} // namespace cr
