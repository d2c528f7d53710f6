// LBVH over triangles (shared by the plain and the robust integer terms).
//
// Build (once per mesh): Morton codes of the triangle centroids -> LSD radix sort (8-bit digits; per-tile LDS
// histograms, stable in-tile ranking with 64-lane ballots) -> Karras (2012) binary radix tree over the sorted codes
// (ties broken by sorted position) -> tree depth. Refit (per box set): bottom-up with one atomic arrival counter per
// internal node; the second child to arrive writes the parent box.
//
// Node encoding: internal nodes 0..n-2 (root 0), leaves 0..n-1 in Morton order; a child index c >= 0 is internal
// node c, c < 0 is leaf ~c.
#pragma once

#include "agwn_internal.hpp"

namespace agwn
{
struct lbvh_topology
{
    int64_t n = 0;          // triangles (leaves)
    int32_t* order = nullptr;   // [n] sorted position -> original triangle index
    int32_t* child = nullptr;   // [2*(n-1)] children of internal nodes
    int32_t* parent = nullptr;  // [n-1 + n]: parent of internal i at [i], of leaf k at [n-1+k]; root's parent -1
    int8_t* big = nullptr;      // [n-1]: 1 when internal i's right subtree has more leaves than its left
    uint32_t* csz = nullptr;    // [n-1]: leaves under internal i's children, min(.., 255) each: left | right << 8
    int max_depth = 0;          // edges from the root to the deepest leaf
};

// Device-owned topology (freed with hipFree by release()).
int lbvh_build_topology(float const* centroids, int64_t n, lbvh_topology& topo, hipStream_t s);
void lbvh_release(lbvh_topology& topo);

// Traversal node: the two child boxes and child indices, which child's subtree has more leaves (big = 1: c1), and
// the leaf counts of both subtrees (sz: left | right << 8, each clamped to 255; in the padding of the 64-byte node).
// Traversals that hit both internal children descend into the smaller subtree and push the larger one: every stack
// entry then marks a step into a subtree of at most half its parent's leaves, so a traversal never holds more than
// log2(n) entries (kMaxStack). B = float or double box coordinates.
//
// Leaf buckets: in Karras order an internal child's subtree is a run of consecutive leaves -- the left child g spans
// leaves [g - size + 1, g], the right child g + 1 spans [g + 1, g + size] -- so a traversal may take any internal child
// of at most L leaves as one bucket and test its leaves directly (bucket_span), without visiting the nodes below.
template <class B>
struct alignas(sizeof(B) == 4 ? 64 : 128) bvh_node
{
    B lo0[3], hi0[3], lo1[3], hi1[3];
    int32_t c0, c1;
    int32_t big;
    uint32_t sz;
};
static_assert(sizeof(bvh_node<float>) == 64 && sizeof(bvh_node<double>) == 128, "node layout");

// The leaves [first, first + count) a child c of a node stands for when it is treated as a leaf: a leaf (c < 0,
// count 1), or with Buckets an internal child whose subtree holds at most L leaves (size = its count from sz).
// Returns false for a child to descend into.
template <bool Buckets>
__device__ __forceinline__ bool bucket_span(int32_t c, uint32_t size, bool right, int L, int32_t& first,
                                            int32_t& count)
{
    if (c < 0)
    {
        first = ~c;
        count = 1;
        return true;
    }
    if (Buckets && int(size) <= L)
    {
        first = right ? c : c - int32_t(size) + 1;
        count = int32_t(size);
        return true;
    }
    return false;
}
constexpr int kMaxBucket = 16; // largest leaf bucket (below the 255 clamp of sz)

// Bottom-up refit. leaf_lo/leaf_hi: [n][3] per sorted leaf (an empty box has lo > hi). Writes nodes[n-1].
// scratch_flags: n-1 ints (zeroed here).
template <class B>
int lbvh_refit(lbvh_topology const& topo, B const* leaf_lo, B const* leaf_hi, bvh_node<B>* nodes, B* node_lo,
               B* node_hi, int32_t* scratch_flags, hipStream_t s);

// traversal stack entries (LDS, per lane): > log2 of the largest supported leaf count (lbvh_build_topology checks
// n < 2^27); 28 x 4 B x 128 lanes = 14 KiB per traversal block
constexpr int kMaxStack = 28;
} // namespace agwn
