// rt_math_host.cpp - the product's intersection header (vierkant/rt_math.hpp) compiled for the host, with C
// entry points, so CPU tests can check its host build against the golden vectors and the oracle.
#include <cstdint>

#include "vierkant/rt_math.hpp"

using namespace vierkant::rt;

extern "C" {

uint32_t host_tri_box_overlap(const float *tri9, const float *box6) { return tri_box_overlap(tri9, box6, box6 + 3); }

uint32_t host_tri_tri_overlap(const float *a9, const float *b9) { return tri_tri_overlap(a9, b9); }

//! both Moller-Trumbore forms: returns 1 on INTERSECT, writes t/u/v; `fast` selects mt_fast
int host_mt(const float *tri9, const float *o, const float *d, int fast, float *tuv)
{
    vec3 v0{tri9[0], tri9[1], tri9[2]}, v1{tri9[3], tri9[4], tri9[5]}, v2{tri9[6], tri9[7], tri9[8]};
    vec3 oo{o[0], o[1], o[2]}, dd{d[0], d[1], d[2]};
    float t = 0.f, u = 0.f, v = 0.f;
    bool hit = fast ? mt_fast(oo, dd, v0, v1 - v0, v2 - v0, t, u, v) : mt_reference(oo, dd, v0, v1 - v0, v2 - v0, t, u, v);
    tuv[0] = t, tuv[1] = u, tuv[2] = v;
    return hit ? 1 : 0;
}

}
