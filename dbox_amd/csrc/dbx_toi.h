// dbx_toi.h -- continuous collision for the step kernels: b2Distance (GJK with a
// simplex cache, src/dbox/collision/b2distance.d:30-707) and b2TimeOfImpact
// (conservative advancement on separating axes, b2timeofimpact.d:67-530), one
// shape pair per lane.
//
// The simplex keeps its three vertices as named members and every access by a
// run-time index goes through a select, so the whole state stays in registers.
// Floating-point operations follow the reference's order; the library is built
// with -ffp-contract=off.
#pragma once

#include "dbx_geom.h"

namespace dbx {

constexpr int kToiUnknown = 0, kToiFailed = 1, kToiOverlapped = 2, kToiTouching = 3, kToiSeparated = 4;

// b2DistanceProxy (b2distance.d:30-140): the shape's vertices (circle centre,
// edge ends, polygon hull) and its radius.  No chain shapes.
struct DProxy {
    const float2 *__restrict__ v;
    int count;
    float radius;
};
__device__ __forceinline__ DProxy make_proxy(const Shape *s) {
    DProxy p;
    p.v = s->type == kCircle ? &s->c : s->v; // edge: v[0] = vertex1, v[1] = vertex2
    p.count = s->type == kCircle ? 1 : (s->type == kEdge ? 2 : s->count);
    p.radius = s->radius;
    return p;
}
// The same proxy with its vertices copied to buf (kMaxPolygonVertices entries,
// LDS), so that the many support and evaluate calls of one TOI read them from
// LDS instead of global memory; the copies are independent loads.
__device__ __forceinline__ DProxy make_proxy_in(const Shape *s, float2 *buf) {
    DProxy p = make_proxy(s);
#pragma unroll
    for (int i = 0; i < kMaxPolygonVertices; ++i)
        if (i < p.count) buf[i] = p.v[i];
    p.v = buf;
    return p;
}
// GetSupport (b2distance.d:94-111): first vertex of largest dot product.
__device__ __forceinline__ int support(const DProxy &p, float2 d) {
    int best = 0;
    float bestValue = dot(p.v[0], d);
    for (int i = 1; i < p.count; ++i) {
        const float value = dot(p.v[i], d);
        if (value > bestValue) {
            best = i;
            bestValue = value;
        }
    }
    return best;
}

// b2Sweep (b2math.d:552-593): c0 and a0 at alpha0, c and a at the step's end.
struct Sweep {
    float2 lc, c0, c;
    float a0, a, alpha0;
};
// GetTransform (b2math.d:556-564)
__device__ __forceinline__ Xf sweep_xf(const Sweep &s, float beta) {
    Xf xf;
    xf.p = add(scl(1.0f - beta, s.c0), scl(beta, s.c));
    const float angle = (1.0f - beta) * s.a0 + beta * s.a;
    xf.q = rot(angle);
    xf.p = sub(xf.p, mul(xf.q, s.lc));
    return xf;
}
// Advance (b2math.d:568-575)
__device__ __forceinline__ void sweep_advance(Sweep &s, float alpha) {
    const float beta = (alpha - s.alpha0) / (1.0f - s.alpha0);
    s.c0 = add(s.c0, scl(beta, sub(s.c, s.c0)));
    s.a0 += beta * (s.a - s.a0);
    s.alpha0 = alpha;
}
// Normalize (b2math.d:578-584)
__device__ __forceinline__ void sweep_normalize(Sweep &s) {
    const float twoPi = 2.0f * kPi;
    const float d = twoPi * floorf(s.a0 / twoPi);
    s.a0 -= d;
    s.a -= d;
}

struct SimplexCache {
    float metric;
    int count;
    int iA0, iA1, iA2, iB0, iB1, iB2;
};
struct SVert {
    float2 wA, wB, w;
    float a;
    int iA, iB;
};
struct Simplex {
    SVert v1, v2, v3;
    int count;
};

__device__ __forceinline__ SVert make_svert(const DProxy &pA, const Xf &xfA, int iA, const DProxy &pB, const Xf &xfB,
                                            int iB) {
    SVert s;
    s.iA = iA;
    s.iB = iB;
    s.wA = mul(xfA, pA.v[iA]);
    s.wB = mul(xfB, pB.v[iB]);
    s.w = sub(s.wB, s.wA);
    s.a = 0.0f;
    return s;
}
// GetMetric (b2distance.d:710-728)
__device__ __forceinline__ float simplex_metric(const Simplex &s) {
    if (s.count == 1) return 0.0f;
    if (s.count == 2) {
        const float2 d = sub(s.v1.w, s.v2.w);
        return sqrtf(d.x * d.x + d.y * d.y);
    }
    return crs(sub(s.v2.w, s.v1.w), sub(s.v3.w, s.v1.w));
}
// ReadCache (b2distance.d:574-624)
__device__ __forceinline__ void read_cache(Simplex &s, const SimplexCache &c, const DProxy &pA, const Xf &xfA,
                                           const DProxy &pB, const Xf &xfB) {
    s.count = c.count;
    if (c.count > 0) s.v1 = make_svert(pA, xfA, c.iA0, pB, xfB, c.iB0);
    if (c.count > 1) s.v2 = make_svert(pA, xfA, c.iA1, pB, xfB, c.iB1);
    if (c.count > 2) s.v3 = make_svert(pA, xfA, c.iA2, pB, xfB, c.iB2);
    if (s.count > 1) {
        const float metric1 = c.metric;
        const float metric2 = simplex_metric(s);
        if (metric2 < 0.5f * metric1 || 2.0f * metric1 < metric2 || metric2 < kEps) s.count = 0;
    }
    if (s.count == 0) {
        s.v1 = make_svert(pA, xfA, 0, pB, xfB, 0);
        s.v1.a = 1.0f;
        s.count = 1;
    }
}
// WriteCache (b2distance.d:626-637)
__device__ __forceinline__ void write_cache(const Simplex &s, SimplexCache &c) {
    c.metric = simplex_metric(s);
    c.count = s.count;
    c.iA0 = s.v1.iA;
    c.iB0 = s.v1.iB;
    c.iA1 = s.v2.iA;
    c.iB1 = s.v2.iB;
    c.iA2 = s.v3.iA;
    c.iB2 = s.v3.iB;
}
// Solve2 (b2distance.d:373-406)
__device__ __forceinline__ void solve2(Simplex &s) {
    const float2 w1 = s.v1.w, w2 = s.v2.w;
    const float2 e12 = sub(w2, w1);
    const float d12_2 = -dot(w1, e12);
    if (d12_2 <= 0.0f) {
        s.v1.a = 1.0f;
        s.count = 1;
        return;
    }
    const float d12_1 = dot(w2, e12);
    if (d12_1 <= 0.0f) {
        s.v2.a = 1.0f;
        s.count = 1;
        s.v1 = s.v2;
        return;
    }
    const float inv_d12 = 1.0f / (d12_1 + d12_2);
    s.v1.a = d12_1 * inv_d12;
    s.v2.a = d12_2 * inv_d12;
    s.count = 2;
}
// Solve3 (b2distance.d:413-571)
__device__ __forceinline__ void solve3(Simplex &s) {
    const float2 w1 = s.v1.w, w2 = s.v2.w, w3 = s.v3.w;
    const float2 e12 = sub(w2, w1);
    const float w1e12 = dot(w1, e12), w2e12 = dot(w2, e12);
    const float d12_1 = w2e12, d12_2 = -w1e12;
    const float2 e13 = sub(w3, w1);
    const float w1e13 = dot(w1, e13), w3e13 = dot(w3, e13);
    const float d13_1 = w3e13, d13_2 = -w1e13;
    const float2 e23 = sub(w3, w2);
    const float w2e23 = dot(w2, e23), w3e23 = dot(w3, e23);
    const float d23_1 = w3e23, d23_2 = -w2e23;
    const float n123 = crs(e12, e13);
    const float d123_1 = n123 * crs(w2, w3);
    const float d123_2 = n123 * crs(w3, w1);
    const float d123_3 = n123 * crs(w1, w2);
    if (d12_2 <= 0.0f && d13_2 <= 0.0f) {
        s.v1.a = 1.0f;
        s.count = 1;
        return;
    }
    if (d12_1 > 0.0f && d12_2 > 0.0f && d123_3 <= 0.0f) {
        const float inv_d12 = 1.0f / (d12_1 + d12_2);
        s.v1.a = d12_1 * inv_d12;
        s.v2.a = d12_2 * inv_d12;
        s.count = 2;
        return;
    }
    if (d13_1 > 0.0f && d13_2 > 0.0f && d123_2 <= 0.0f) {
        const float inv_d13 = 1.0f / (d13_1 + d13_2);
        s.v1.a = d13_1 * inv_d13;
        s.v3.a = d13_2 * inv_d13;
        s.count = 2;
        s.v2 = s.v3;
        return;
    }
    if (d12_1 <= 0.0f && d23_2 <= 0.0f) {
        s.v2.a = 1.0f;
        s.count = 1;
        s.v1 = s.v2;
        return;
    }
    if (d13_1 <= 0.0f && d23_1 <= 0.0f) {
        s.v3.a = 1.0f;
        s.count = 1;
        s.v1 = s.v3;
        return;
    }
    if (d23_1 > 0.0f && d23_2 > 0.0f && d123_1 <= 0.0f) {
        const float inv_d23 = 1.0f / (d23_1 + d23_2);
        s.v2.a = d23_1 * inv_d23;
        s.v3.a = d23_2 * inv_d23;
        s.count = 2;
        s.v1 = s.v3;
        return;
    }
    const float inv_d123 = 1.0f / (d123_1 + d123_2 + d123_3);
    s.v1.a = d123_1 * inv_d123;
    s.v2.a = d123_2 * inv_d123;
    s.v3.a = d123_3 * inv_d123;
    s.count = 3;
}

// b2Distance (b2distance.d:185-345) without radii: the distance of the two
// shapes' cores and the simplex left in the cache.
__device__ float gjk_distance(SimplexCache &cache, const DProxy &pA, const Xf &xfA, const DProxy &pB, const Xf &xfB) {
    Simplex s;
    read_cache(s, cache, pA, xfA, pB, xfB);
    int sA0 = 0, sA1 = 0, sA2 = 0, sB0 = 0, sB1 = 0, sB2 = 0;
    for (int iter = 0; iter < 20;) {
        const int saveCount = s.count;
        sA0 = s.v1.iA;
        sB0 = s.v1.iB;
        sA1 = s.v2.iA;
        sB1 = s.v2.iB;
        sA2 = s.v3.iA;
        sB2 = s.v3.iB;
        if (s.count == 2) solve2(s);
        else if (s.count == 3) solve3(s);
        if (s.count == 3) break;
        // GetSearchDirection (b2distance.d:639-662)
        float2 d;
        if (s.count == 1) {
            d = neg(s.v1.w);
        } else {
            const float2 e12 = sub(s.v2.w, s.v1.w);
            const float sgn = crs(e12, neg(s.v1.w));
            d = sgn > 0.0f ? crs(1.0f, e12) : crs(e12, 1.0f);
        }
        if (lensq(d) < kEps * kEps) break;
        const int iA = support(pA, mulT(xfA.q, neg(d)));
        const int iB = support(pB, mulT(xfB.q, d));
        const SVert nv = make_svert(pA, xfA, iA, pB, xfB, iB);
        ++iter;
        const bool dup = (saveCount > 0 && iA == sA0 && iB == sB0) || (saveCount > 1 && iA == sA1 && iB == sB1) ||
                         (saveCount > 2 && iA == sA2 && iB == sB2);
        if (dup) break;
        if (s.count == 1) s.v2 = nv;
        else s.v3 = nv;
        ++s.count;
    }
    // GetWitnessPoints (b2distance.d:686-708)
    float2 wa, wb;
    if (s.count == 1) {
        wa = s.v1.wA;
        wb = s.v1.wB;
    } else if (s.count == 2) {
        wa = add(scl(s.v1.a, s.v1.wA), scl(s.v2.a, s.v2.wA));
        wb = add(scl(s.v1.a, s.v1.wB), scl(s.v2.a, s.v2.wB));
    } else {
        wa = add(add(scl(s.v1.a, s.v1.wA), scl(s.v2.a, s.v2.wA)), scl(s.v3.a, s.v3.wA));
        wb = wa;
    }
    write_cache(s, cache);
    const float2 dd = sub(wa, wb);
    return sqrtf(dd.x * dd.x + dd.y * dd.y);
}

// b2TestOverlap of two shapes (b2collision.d:449-468): b2Distance with radii from
// an empty cache (b2distance.d:327-344); overlapping below 10 epsilon.
__device__ __noinline__ bool test_overlap(const Shape *a, const Xf &xfA, const Shape *b, const Xf &xfB) {
    SimplexCache cache;
    cache.count = 0;
    const DProxy pA = make_proxy(a), pB = make_proxy(b);
    const float d = gjk_distance(cache, pA, xfA, pB, xfB);
    const float r = pA.radius + pB.radius;
    const float dist = (d > r && d > kEps) ? d - r : 0.0f;
    return dist < 10.0f * kEps;
}

// b2SeparationFunction (b2timeofimpact.d:297-530)
struct SepFn {
    Sweep sA, sB;
    int type; // 0 points, 1 faceA, 2 faceB
    float2 lp, axis;
};
// Initialize (b2timeofimpact.d:331-410)
__device__ void sep_init(SepFn &f, const SimplexCache &cache, const DProxy &pA, const Sweep &sA, const DProxy &pB,
                         const Sweep &sB, float t1) {
    f.sA = sA;
    f.sB = sB;
    const Xf xfA = sweep_xf(sA, t1), xfB = sweep_xf(sB, t1);
    if (cache.count == 1) {
        f.type = 0;
        const float2 pointA = mul(xfA, pA.v[cache.iA0]);
        const float2 pointB = mul(xfB, pB.v[cache.iB0]);
        f.axis = sub(pointB, pointA);
        normalize(f.axis);
        f.lp = v2(0.0f, 0.0f);
    } else if (cache.iA0 == cache.iA1) {
        f.type = 2;
        const float2 b1 = pB.v[cache.iB0], b2 = pB.v[cache.iB1];
        f.axis = crs(sub(b2, b1), 1.0f);
        normalize(f.axis);
        const float2 normal = mul(xfB.q, f.axis);
        f.lp = scl(0.5f, add(b1, b2));
        const float2 pointB = mul(xfB, f.lp);
        const float2 pointA = mul(xfA, pA.v[cache.iA0]);
        const float s = dot(sub(pointA, pointB), normal);
        if (s < 0.0f) f.axis = neg(f.axis);
    } else {
        f.type = 1;
        const float2 a1 = pA.v[cache.iA0], a2 = pA.v[cache.iA1];
        f.axis = crs(sub(a2, a1), 1.0f);
        normalize(f.axis);
        const float2 normal = mul(xfA.q, f.axis);
        f.lp = scl(0.5f, add(a1, a2));
        const float2 pointA = mul(xfA, f.lp);
        const float2 pointB = mul(xfB, pB.v[cache.iB0]);
        const float s = dot(sub(pointB, pointA), normal);
        if (s < 0.0f) f.axis = neg(f.axis);
    }
}
// FindMinSeparation (b2timeofimpact.d:413-470)
__device__ float sep_find_min(const SepFn &f, const DProxy &pA, const DProxy &pB, int &iA, int &iB, float t) {
    const Xf xfA = sweep_xf(f.sA, t), xfB = sweep_xf(f.sB, t);
    if (f.type == 0) {
        iA = support(pA, mulT(xfA.q, f.axis));
        iB = support(pB, mulT(xfB.q, neg(f.axis)));
        const float2 pointA = mul(xfA, pA.v[iA]);
        const float2 pointB = mul(xfB, pB.v[iB]);
        return dot(sub(pointB, pointA), f.axis);
    }
    if (f.type == 1) {
        const float2 normal = mul(xfA.q, f.axis);
        const float2 pointA = mul(xfA, f.lp);
        iA = -1;
        iB = support(pB, mulT(xfB.q, neg(normal)));
        const float2 pointB = mul(xfB, pB.v[iB]);
        return dot(sub(pointB, pointA), normal);
    }
    const float2 normal = mul(xfB.q, f.axis);
    const float2 pointB = mul(xfB, f.lp);
    iB = -1;
    iA = support(pA, mulT(xfA.q, neg(normal)));
    const float2 pointA = mul(xfA, pA.v[iA]);
    return dot(sub(pointA, pointB), normal);
}
// Evaluate (b2timeofimpact.d:473-518)
__device__ float sep_eval(const SepFn &f, const DProxy &pA, const DProxy &pB, int iA, int iB, float t) {
    const Xf xfA = sweep_xf(f.sA, t), xfB = sweep_xf(f.sB, t);
    if (f.type == 0) {
        const float2 pointA = mul(xfA, pA.v[iA]);
        const float2 pointB = mul(xfB, pB.v[iB]);
        return dot(sub(pointB, pointA), f.axis);
    }
    if (f.type == 1) {
        const float2 normal = mul(xfA.q, f.axis);
        const float2 pointA = mul(xfA, f.lp);
        const float2 pointB = mul(xfB, pB.v[iB]);
        return dot(sub(pointB, pointA), normal);
    }
    const float2 normal = mul(xfB.q, f.axis);
    const float2 pointB = mul(xfB, f.lp);
    const float2 pointA = mul(xfA, pA.v[iA]);
    return dot(sub(pointA, pointB), normal);
}

// b2TimeOfImpact (b2timeofimpact.d:67-295) over [0, tMax]; returns the output
// state and writes the time.  Every loop is bounded as in the reference (20
// outer iterations, b2_maxPolygonVertices push-backs, 50 root-finder steps).
__device__ int time_of_impact(float &tOut, const DProxy &pA, const DProxy &pB, Sweep sA, Sweep sB, float tMax) {
    int state = kToiUnknown;
    tOut = tMax;
    sweep_normalize(sA);
    sweep_normalize(sB);
    const float totalRadius = pA.radius + pB.radius;
    const float target = fmaxr(kLinearSlop, totalRadius - 3.0f * kLinearSlop);
    const float tolerance = 0.25f * kLinearSlop;
    float t1 = 0.0f;
    SimplexCache cache;
    cache.count = 0;
    cache.metric = 0.0f;
    cache.iA0 = cache.iA1 = cache.iA2 = cache.iB0 = cache.iB1 = cache.iB2 = 0;
    for (int iter = 0;;) {
        const Xf xfA = sweep_xf(sA, t1), xfB = sweep_xf(sB, t1);
        const float distance = gjk_distance(cache, pA, xfA, pB, xfB);
        if (distance <= 0.0f) {
            state = kToiOverlapped;
            tOut = 0.0f;
            break;
        }
        if (distance < target + tolerance) {
            state = kToiTouching;
            tOut = t1;
            break;
        }
        SepFn fcn;
        sep_init(fcn, cache, pA, sA, pB, sB, t1);
        bool done = false;
        float t2 = tMax;
        for (int pushBackIter = 0;;) {
            int iA, iB;
            float s2 = sep_find_min(fcn, pA, pB, iA, iB, t2);
            if (s2 > target + tolerance) {
                state = kToiSeparated;
                tOut = tMax;
                done = true;
                break;
            }
            if (s2 > target - tolerance) {
                t1 = t2;
                break;
            }
            float s1 = sep_eval(fcn, pA, pB, iA, iB, t1);
            if (s1 < target - tolerance) {
                state = kToiFailed;
                tOut = t1;
                done = true;
                break;
            }
            if (s1 <= target + tolerance) {
                state = kToiTouching;
                tOut = t1;
                done = true;
                break;
            }
            float a1 = t1, a2 = t2;
            for (int rootIterCount = 0;;) {
                float t;
                if (rootIterCount & 1) t = a1 + (target - s1) * (a2 - a1) / (s2 - s1);
                else t = 0.5f * (a1 + a2);
                ++rootIterCount;
                const float s = sep_eval(fcn, pA, pB, iA, iB, t);
                if (fabsr(s - target) < tolerance) {
                    t2 = t;
                    break;
                }
                if (s > target) {
                    a1 = t;
                    s1 = s;
                } else {
                    a2 = t;
                    s2 = s;
                }
                if (rootIterCount == 50) break;
            }
            ++pushBackIter;
            if (pushBackIter == kMaxPolygonVertices) break;
        }
        ++iter;
        if (done) break;
        if (iter == 20) {
            state = kToiFailed;
            tOut = t1;
            break;
        }
    }
    return state;
}

} // namespace dbx
