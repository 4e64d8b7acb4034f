// ref_driver.cpp - C entry points over the reference's own src/triangle_intersection.h.
//
// TEST INFRASTRUCTURE ONLY. The header is compiled where it lies under the reference checkout
// (include path given by oracle/Makefile); nothing of it is copied. The wrappers mirror
// intersect(const Triangle&, const AABB&) and intersect(const Triangle&, const Triangle&) of
// src/intersection.cpp:364-378: box center = (max + min) / 2.f, half extents = (max - min) / 2.f.
#include <cmath>
#include <cstdint>

#include "triangle_intersection.h"

extern "C" {

uint32_t ref_tri_box_overlap(const float *tri9, const float *box6)
{
    float center[3], half[3];
    for(int a = 0; a < 3; ++a)
    {
        center[a] = (box6[3 + a] + box6[a]) / 2.f;
        half[a] = (box6[3 + a] - box6[a]) / 2.f;
    }
    return (uint32_t)tri_box_overlap(center, half, reinterpret_cast<const float(*)[3]>(tri9));
}

uint32_t ref_tri_tri_overlap(const float *a9, const float *b9)
{
    return (uint32_t)tri_tri_overlap_test_3d(a9, a9 + 3, a9 + 6, b9, b9 + 3, b9 + 6);
}

}
