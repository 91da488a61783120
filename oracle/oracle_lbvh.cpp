/*
 * ORACLE — test infrastructure only.
 *
 * CPU restatement of the photon-beam pipeline of the reference
 * (RenderTechniquePPB): beam production (shaders/PPB_PT.comp), Morton codes
 * (PPB_PT.comp:317-333), the stable LSD radix sort
 * (PPB_RadixSort_{LocalSort,PrefixSum,GlobalSort}.comp), the Karras radix
 * tree (PPB_GenerateHierarchy.comp), the AABB refit (PPB_CalculateAABB.comp)
 * and the beam gather (PPB_PE.comp).  The reference's LBVH shaders are never
 * dispatched (RenderTechniquePPB.cpp:436-519) and carry defects; the method
 * is restated with those defects corrected: 30-bit Morton codes, a stable
 * LSD radix sort, an exact integer clz with an index tie-break, findSplit on
 * (min, max), innerCount = n - 1, leaves that index the sorted beams, and
 * unions over bounds[1].
 */
#include <math.h>
#include <stdint.h>
#include <string.h>
#include <algorithm>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle_core.h"

using namespace orc;

extern "C" {

/* ------------------------------------------------------------------ Morton
 * expandBits — PPB_PT.comp:317-323 is the 8-bit form of Karras's 10-bit
 * expansion ("Thinking Parallel III"); the 10-bit form adds the first mask. */
uint32_t or_expand_bits10(uint32_t v)
{
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

static inline uint32_t expand_bits8(uint32_t v)
{
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

/* getMortonCode — PPB_PT.comp:326-333 (24-bit, absolute /256 quantisation). */
uint32_t or_morton24_ref(const float pos[3], const float b0[3])
{
    float rx = pos[0] - b0[0], ry = pos[1] - b0[1], rz = pos[2] - b0[2];
    uint32_t x = (uint32_t)fminf(fmaxf(rx / 256.0f, 0.0f), 255.0f);
    uint32_t y = (uint32_t)fminf(fmaxf(ry / 256.0f, 0.0f), 255.0f);
    uint32_t z = (uint32_t)fminf(fmaxf(rz / 256.0f, 0.0f), 255.0f);
    return (expand_bits8(x) << 2) + (expand_bits8(y) << 1) + expand_bits8(z);
}

/* 30-bit Morton code of a point normalised to the box [b0,b1]. */
uint32_t or_morton30(const float c[3], const float b0[3], const float b1[3])
{
    uint32_t q[3];
    for (int k = 0; k < 3; k++) {
        float n = (c[k] - b0[k]) / (b1[k] - b0[k]);
        q[k] = (uint32_t)fminf(fmaxf(n * 1024.0f, 0.0f), 1023.0f);
    }
    return (or_expand_bits10(q[0]) << 2) + (or_expand_bits10(q[1]) << 1) + or_expand_bits10(q[2]);
}

/* Morton code of each beam's centroid (startPos + endPos) / 2 (PPB_PT.comp:387). */
void or_beam_morton30(const cr_photon_beam* beams, uint32_t n, const float b0[3], const float b1[3], uint32_t* out)
{
    for (uint32_t i = 0; i < n; i++) {
        float c[3];
        for (int k = 0; k < 3; k++) c[k] = (beams[i].startPos[k] + beams[i].endPos[k]) / 2.0f;
        out[i] = or_morton30(c, b0, b1);
    }
}

/* ------------------------------------------------------------------- sort
 * Stable LSD radix sort of (key, index) with 8-bit digits, 4 passes.  The
 * reference intends 4-bit digits through LocalSort/PrefixSum/GlobalSort
 * (RenderTechniquePPB.cpp:464-491); any stable LSD sort yields the unique
 * permutation "by key, ties by input index". */
void or_radix_sort_pairs(const uint32_t* keys, uint32_t n, uint32_t* keys_out, uint32_t* perm_out)
{
    std::vector<uint32_t> k0(keys, keys + n), k1(n), p0(n), p1(n);
    for (uint32_t i = 0; i < n; i++) p0[i] = i;
    for (int pass = 0; pass < 4; pass++) {
        uint32_t shift = 8u * (uint32_t)pass;
        uint32_t count[257] = {0};
        for (uint32_t i = 0; i < n; i++) count[((k0[i] >> shift) & 0xffu) + 1]++;
        for (int d = 0; d < 256; d++) count[d + 1] += count[d];
        for (uint32_t i = 0; i < n; i++) {
            uint32_t d = (k0[i] >> shift) & 0xffu;
            uint32_t dst = count[d]++;
            k1[dst] = k0[i];
            p1[dst] = p0[i];
        }
        k0.swap(k1);
        p0.swap(p1);
    }
    memcpy(keys_out, k0.data(), sizeof(uint32_t) * n);
    memcpy(perm_out, p0.data(), sizeof(uint32_t) * n);
}

} // extern "C"

namespace {

inline int clz32(uint32_t v) { return v ? __builtin_clz(v) : 32; }

/* delta(i,j) of Karras 2012 with the index tie-break for equal keys
 * (corrected: exact integer clz, duplicates split by index). */
inline int delta(const uint32_t* k, int64_t n, int64_t i, int64_t j)
{
    if (j < 0 || j > n - 1) return -1;
    uint32_t a = k[i], b = k[j];
    if (a == b) return 32 + clz32((uint32_t)i ^ (uint32_t)j);
    return clz32(a ^ b);
}

const uint32_t NONE = 0xFFFFFFFFu;

void init_nodes(cr_tree_node* nodes, uint32_t n)
{
    uint32_t total = n ? 2 * n - 1 : 0;
    memset(nodes, 0, sizeof(cr_tree_node) * total);
    for (uint32_t i = 0; i < total; i++) {
        nodes[i].parent = NONE;
        nodes[i].isLeaf = (i >= n - 1) ? 1u : 0u;
    }
    for (uint32_t i = 0; i < n; i++) {
        nodes[n - 1 + i].left = i;    /* leaves index the sorted beams */
        nodes[n - 1 + i].right = NONE;
    }
}

/* Karras split search over [first, last] using delta (ties by index). */
int64_t find_split(const uint32_t* k, int64_t n, int64_t first, int64_t last)
{
    int dnode = delta(k, n, first, last);
    int64_t s = first;
    int64_t step = last - first;
    do {
        step = (step + 1) >> 1;
        int64_t ns = s + step;
        if (ns < last && delta(k, n, first, ns) > dnode) s = ns;
    } while (step > 1);
    return s;
}

void topdown(const uint32_t* k, int64_t n, int64_t first, int64_t last, uint32_t idx, cr_tree_node* nodes)
{
    int64_t g = find_split(k, n, first, last);
    uint32_t inner = (uint32_t)(n - 1);
    uint32_t left = (first == g) ? inner + (uint32_t)g : (uint32_t)g;
    uint32_t right = (last == g + 1) ? inner + (uint32_t)(g + 1) : (uint32_t)(g + 1);
    nodes[idx].left = left;
    nodes[idx].right = right;
    nodes[left].parent = idx;
    nodes[right].parent = idx;
    if (first != g) topdown(k, n, first, g, (uint32_t)g, nodes);
    if (last != g + 1) topdown(k, n, g + 1, last, (uint32_t)(g + 1), nodes);
}

} // namespace

extern "C" {

/* PPB_GenerateHierarchy.comp main (:122-169), per internal node i in
 * [0, n-2]: direction, range and split of Karras 2012 "Maximizing Parallelism
 * in the Construction of BVHs, Octrees, and k-d Trees", Figure 4, with
 * findSplit called on (min, max) of the range and innerCount = n-1.
 * Node layout as the reference: internal nodes [0, n-2], leaves at
 * innerCount + i, root 0 (parent NONE). */
void or_karras_build(const uint32_t* k, uint32_t n32, cr_tree_node* nodes)
{
    if (n32 == 0) return;
    init_nodes(nodes, n32);
    int64_t n = n32;
    uint32_t inner = n32 - 1;
    for (int64_t i = 0; i < n - 1; i++) {
        int d = (delta(k, n, i, i + 1) - delta(k, n, i, i - 1)) > 0 ? 1 : -1;
        int dmin = delta(k, n, i, i - d);
        int64_t lmax = 2;
        while (delta(k, n, i, i + lmax * d) > dmin) lmax *= 2;
        int64_t l = 0;
        for (int64_t t = lmax / 2; t >= 1; t /= 2)
            if (delta(k, n, i, i + (l + t) * d) > dmin) l += t;
        int64_t j = i + l * d;
        int64_t first = std::min(i, j), last = std::max(i, j);
        int64_t g = find_split(k, n, first, last);
        uint32_t left = (first == g) ? inner + (uint32_t)g : (uint32_t)g;
        uint32_t right = (last == g + 1) ? inner + (uint32_t)(g + 1) : (uint32_t)(g + 1);
        nodes[i].left = left;
        nodes[i].right = right;
        nodes[left].parent = (uint32_t)i;
        nodes[right].parent = (uint32_t)i;
    }
}

/* Top-down recursive builder over the same sorted keys: split each range at
 * the highest differing bit (ties by index).  It produces the same layout as
 * the parallel form and is used only to check it. */
void or_topdown_build(const uint32_t* k, uint32_t n, cr_tree_node* nodes)
{
    if (n == 0) return;
    init_nodes(nodes, n);
    if (n == 1) return;
    topdown(k, n, 0, n - 1, 0, nodes);
}

/* calculateBeamBounds — PPB_CalculateAABB.comp:64-80 (iq's capped-cylinder
 * box).  A zero-length beam gets the ball of its radius (the reference would
 * produce NaN). */
void or_leaf_bounds(const cr_photon_beam* b, float lo[3], float hi[3])
{
    f3 pa = ld3(b->startPos), pb = ld3(b->endPos);
    f3 a = sub(pb, pa);
    float dd = dot(a, a);
    float ra = b->radius;
    f3 e;
    if (dd > 0.0f) {
        e = mk(ra * sqrtf(1.0f - (a.x * a.x) / dd), ra * sqrtf(1.0f - (a.y * a.y) / dd),
               ra * sqrtf(1.0f - (a.z * a.z) / dd));
    } else {
        e = mk(ra, ra, ra);
    }
    lo[0] = fminf(pa.x - e.x, pb.x - e.x); hi[0] = fmaxf(pa.x + e.x, pb.x + e.x);
    lo[1] = fminf(pa.y - e.y, pb.y - e.y); hi[1] = fmaxf(pa.y + e.y, pb.y + e.y);
    lo[2] = fminf(pa.z - e.z, pb.z - e.z); hi[2] = fmaxf(pa.z + e.z, pb.z + e.z);
}

/* PPB_CalculateAABB.comp main (:100-120): leaf boxes, then the union of the
 * two children (corrected: max over bounds[1]); internal nodes end with
 * processed = 2 (both children arrived), as the atomic rule leaves them. */
void or_refit(const cr_photon_beam* sorted_beams, uint32_t n, cr_tree_node* nodes)
{
    if (n == 0) return;
    uint32_t inner = n - 1;
    for (uint32_t i = 0; i < n; i++) {
        float lo[3], hi[3];
        or_leaf_bounds(&sorted_beams[i], lo, hi);
        cr_tree_node& L = nodes[inner + i];
        for (int k = 0; k < 3; k++) { L.bounds[0][k] = lo[k]; L.bounds[1][k] = hi[k]; }
        L.bounds[0][3] = 0.0f;
        L.bounds[1][3] = 0.0f;
        L.processed = 0;
    }
    if (n == 1) return;
    /* post-order over internal nodes with an explicit stack */
    std::vector<std::pair<uint32_t, int>> st;
    st.push_back({0u, 0});
    while (!st.empty()) {
        auto& top = st.back();
        uint32_t idx = top.first;
        cr_tree_node& N = nodes[idx];
        if (top.second == 0) {
            top.second = 1;
            if (!nodes[N.left].isLeaf) st.push_back({N.left, 0});
            if (!nodes[N.right].isLeaf) st.push_back({N.right, 0});
            continue;
        }
        const cr_tree_node& A = nodes[N.left];
        const cr_tree_node& B = nodes[N.right];
        for (int k = 0; k < 4; k++) {
            N.bounds[0][k] = fminf(A.bounds[0][k], B.bounds[0][k]);
            N.bounds[1][k] = fmaxf(A.bounds[1][k], B.bounds[1][k]);
        }
        N.processed = 2;
        st.pop_back();
    }
}

/* Whole build: Morton30 of centroids in the cloud box, stable sort, Karras
 * hierarchy over the sorted codes, refit over the beams in sorted order. */
void or_lbvh_build(const cr_photon_beam* beams, uint32_t n, const float b0[3], const float b1[3],
                   uint32_t* morton_sorted, uint32_t* perm, cr_tree_node* nodes)
{
    if (n == 0) return;
    std::vector<uint32_t> codes(n);
    or_beam_morton30(beams, n, b0, b1, codes.data());
    or_radix_sort_pairs(codes.data(), n, morton_sorted, perm);
    or_karras_build(morton_sorted, n, nodes);
    std::vector<cr_photon_beam> sorted(n);
    for (uint32_t i = 0; i < n; i++) {
        sorted[i] = beams[perm[i]];
        sorted[i].mortonCode = morton_sorted[i];
    }
    or_refit(sorted.data(), n, nodes);
}

} // extern "C"
