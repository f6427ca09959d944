// Device-visible BVH record layouts (host and device code include this).
//
// Replaces the Embree scene (include/nanogi/rt.hpp:2087-2143).  Triangles are
// stored in leaf order; every child reference is a 32-bit word:
//   >= 0 : index of an internal node record
//   <  0 : leaf, ~ref = (first_leaf_triangle << 4) | (count - 1)
#pragma once

#include <stdint.h>

namespace ngi {

// Binary node, 64 bytes: both child boxes and both child references in one
// record, so one visit is four 16-byte loads.
struct alignas(16) Bvh2Node {
    float c0lox, c0hix, c0loy, c0hiy;
    float c1lox, c1hix, c1loy, c1hiy;
    float c0loz, c0hiz, c1loz, c1hiz;
    int32_t child0, child1, pad0, pad1;
};
static_assert(sizeof(Bvh2Node) == 64, "Bvh2Node must be 64 bytes");

// Wide node, 128 bytes (one L2 line): up to eight children with bounds
// quantised to 8 bits on a per-node power-of-two grid (after Ylitie, Karras,
// Laine, "Efficient Incoherent Ray Traversal on GPUs Through Compressed Wide
// BVHs", HPG 2017), laid out for one 16-byte load per four fields.
//   origin.xyz + exponents packed: child box k spans
//     lo = origin + qlo[k] * 2^e,  hi = origin + qhi[k] * 2^e  (per axis)
//   child[k]: reference as above, or NGI_BVH_EMPTY for an unused slot.
// Slot k of a node holds the child whose centroid lies in octant k of the
// parent box (bit 0: +x, bit 1: +y, bit 2: +z), resolved greedily when two
// children compete, so XOR-ing slot indices with a ray's direction octant
// gives an approximately near-to-far visiting order.
struct alignas(16) Bvh8Node {
    float ox, oy, oz;
    uint8_t ex, ey, ez;              // biased exponents (e + 127)
    uint8_t valid;                   // bit k set if slot k holds a child
    uint8_t qlox[8], qloy[8], qloz[8];
    uint8_t qhix[8], qhiy[8], qhiz[8];
    int32_t child[8];
    uint32_t pad[8];
};
static_assert(sizeof(Bvh8Node) == 128, "Bvh8Node must be 128 bytes");

#define NGI_BVH_EMPTY ((int32_t)0x7FFFFFFF)

static inline int32_t ngi_leaf_ref(uint32_t first, uint32_t count) { return ~(int32_t)((first << 4) | (count - 1)); }

}  // namespace ngi
