"""numpy dtypes of the POD records in include/palz.h (reference layouts)."""
from __future__ import annotations

import numpy as np

# minimum_lbvh::Triangle (minimum_lbvh.h:207-210): 3 x float3, 36 B
TRIANGLE = np.dtype([("vs", "<f4", (3, 3))])
assert TRIANGLE.itemsize == 36

# minimum_lbvh::AABB (minimum_lbvh.h:220-261): 24 B
AABB = np.dtype([("lower", "<f4", 3), ("upper", "<f4", 3)])
assert AABB.itemsize == 24

# minimum_lbvh::InternalNode (minimum_lbvh.h:285-291): 64 B
INTERNAL_NODE = np.dtype([("children", "<u4", 2), ("aabbs", AABB, 2), ("parent", "<u4"),
                          ("context", "<u4")])
assert INTERNAL_NODE.itemsize == 64

# palz_wide_node: the BVH4 view of one internal node, its six boxes by plane:
# plane[axis][lower, upper, lower][slot], 248 B
WIDE_NODE = np.dtype([("child", "<u4", 2), ("grand", "<u4", 4), ("pad", "<u4", 2), ("plane", "<f4", (3, 3, 6))])
assert WIDE_NODE.itemsize == 248

# InternalNormalBound (pk.h:1705-1709): 2 AABBs of shading normals + arrival counter, 52 B
NORMAL_BOUND = np.dtype([("normal_bounds", AABB, 2), ("counter", "<u4")])
assert NORMAL_BOUND.itemsize == 52

# TriangleAttrib (pk.h:74-78): int material + 3 x float3 shading normals, 40 B
TRIANGLE_ATTRIB = np.dtype([("material", "<i4"), ("shading_normals", "<f4", (3, 3))])
assert TRIANGLE_ATTRIB.itemsize == 40

NODE_INVALID = 0x7FFFFFFF
NODE_LEAF_BIT = 0x80000000

MATERIAL_DIFFUSE = 0
MATERIAL_MIRROR = 1
MATERIAL_DIELECTRIC = 2

BUILD_OPTION_USE_NORMAL = 1
BUILD_OPTION_CPU_PARALLEL = 2

# palz_hit (minimum_lbvh::Hit, minimum_lbvh.h:905-911)
HIT = np.dtype([("t", "<f4"), ("u", "<f4"), ("v", "<f4"), ("ng", "<f4", 3), ("tri", "<u4")])
assert HIT.itemsize == 28
QUERY_CLOSEST, QUERY_CLOSEST_STACKFREE, QUERY_ANY = 0, 1, 2
