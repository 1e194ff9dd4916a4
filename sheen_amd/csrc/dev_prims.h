// dev_prims.h — ray/primitive tests of the HIP path.
//
// Each test follows the reference's arithmetic step for step (cited per
// function) because hit distances must agree bit for bit with the CPU
// oracle; only the data layout around them is new.
#pragma once
#include "dev_common.h"

// Reference slab test (bvh.h:578-600) against one primitive's own bounds:
// the gate a primitive passes before its test in the reference's traversal.
SHD bool ref_slab(const DRay& r, v3 inv, const float* mn, const float* mx) {
    float tminX = (mn[0] - r.o.x) * inv.x;
    float tmaxX = (mx[0] - r.o.x) * inv.x;
    if (tminX > tmaxX) { float t = tminX; tminX = tmaxX; tmaxX = t; }
    float tminY = (mn[1] - r.o.y) * inv.y;
    float tmaxY = (mx[1] - r.o.y) * inv.y;
    if (tminY > tmaxY) { float t = tminY; tminY = tmaxY; tmaxY = t; }
    if (tminX > tmaxY || tminY > tmaxX) return false;
    if (tminY > tminX) tminX = tminY;
    if (tmaxY < tmaxX) tmaxX = tmaxY;
    float tminZ = (mn[2] - r.o.z) * inv.z;
    float tmaxZ = (mx[2] - r.o.z) * inv.z;
    if (tminZ > tmaxZ) { float t = tminZ; tminZ = tmaxZ; tmaxZ = t; }
    if (tminX > tmaxZ || tminZ > tmaxX) return false;
    return (tmaxX >= r.tmin) && (tminX <= r.tmax);
}

// Sphere (reference primitives/sphere.h:30-53).  uv: (0,0).
SHD bool hit_sphere(const DRay& r, v3 c, float radius, DHit& h) {
    v3 oc = r.o - c;
    float a = dot(r.d, r.d);
    float b = 2.0f * dot(oc, r.d);
    float cc = dot(oc, oc) - radius * radius;
    float disc = b * b - 4.0f * a * cc;
    if (disc < 0.0f) return false;
    float sq = sqrtf(disc);
    float t = (-b - sq) / (2.0f * a);
    if (t < r.tmin || t > r.tmax) {
        t = (-b + sq) / (2.0f * a);
        if (t < r.tmin || t > r.tmax) return false;
    }
    h.t = t;
    h.point = r.o + t * r.d;
    h.normal = normalize(h.point - c);
    h.tangent = mk3(0, 0, 0);
    h.uvx = 0.0f; h.uvy = 0.0f;
    return true;
}
SHD bool t_sphere(const DRay& r, v3 c, float radius, float& t) {
    v3 oc = r.o - c;
    float a = dot(r.d, r.d);
    float b = 2.0f * dot(oc, r.d);
    float cc = dot(oc, oc) - radius * radius;
    float disc = b * b - 4.0f * a * cc;
    if (disc < 0.0f) return false;
    float sq = sqrtf(disc);
    t = (-b - sq) / (2.0f * a);
    if (t < r.tmin || t > r.tmax) {
        t = (-b + sq) / (2.0f * a);
        if (t < r.tmin || t > r.tmax) return false;
    }
    return true;
}
// reference primitives/sphere.h:8-25
SHD bool sphere_span(const DRay& r, v3 c, float radius, float t0, float t1, float& te, float& tx) {
    v3 oc = r.o - c;
    float a = dot(r.d, r.d);
    float b = 2.0f * dot(oc, r.d);
    float cc = dot(oc, oc) - radius * radius;
    float disc = b * b - 4.0f * a * cc;
    if (disc < 0.0f) return false;
    float sq = sqrtf(disc);
    float ta = (-b - sq) / (2.0f * a);
    float tb = (-b + sq) / (2.0f * a);
    if (tb < t0 || ta > t1) return false;
    te = fmaxf(ta, t0);
    tx = fminf(tb, t1);
    return te <= tx;
}

// Plane (reference primitives/plane.h:6-24).  uv: (x, z) of the hit point.
SHD bool hit_plane(const DRay& r, v3 n, float d, DHit& h) {
    float denom = dot(n, r.d);
    if (fabsf(denom) > 1e-6f) {
        float t = -(dot(r.o, n) + d) / denom;
        if (t >= r.tmin && t <= r.tmax) {
            h.t = t;
            h.point = r.o + t * r.d;
            h.normal = n;
            h.tangent = mk3(0, 0, 0);
            if (dot(h.normal, r.d) > 0.0f) h.normal = -h.normal;
            h.uvx = h.point.x; h.uvy = h.point.z;
            return true;
        }
    }
    return false;
}

// Moller-Trumbore on the leaf copy of the vertices (reference
// primitives/triangle.h:7-46, distance part only).
SHD bool t_triangle(const DRay& r, v3 v0, v3 v1, v3 v2, float& t) {
    v3 e1 = v1 - v0, e2 = v2 - v0;
    v3 pvec = cross(r.d, e2);
    float det = dot(e1, pvec);
    if (fabsf(det) < 1e-8f) return false;
    float inv_det = 1.0f / det;
    v3 tvec = r.o - v0;
    float u = dot(tvec, pvec) * inv_det;
    if (u < 0.0f || u > 1.0f) return false;
    v3 qvec = cross(tvec, e1);
    float v = dot(r.d, qvec) * inv_det;
    if (v < 0.0f || u + v > 1.0f) return false;
    t = dot(e2, qvec) * inv_det;
    if (t < r.tmin || t > r.tmax) return false;
    return true;
}
// Full triangle record (reference primitives/triangle.h:7-46).
SHD bool hit_triangle(const DRay& r, const Triangle& tri, DHit& h) {
    v3 v0 = cv3(tri.v0), v1 = cv3(tri.v1), v2 = cv3(tri.v2);
    v3 e1 = v1 - v0, e2 = v2 - v0;
    v3 pvec = cross(r.d, e2);
    float det = dot(e1, pvec);
    if (fabsf(det) < 1e-8f) return false;
    float inv_det = 1.0f / det;
    v3 tvec = r.o - v0;
    float u = dot(tvec, pvec) * inv_det;
    if (u < 0.0f || u > 1.0f) return false;
    v3 qvec = cross(tvec, e1);
    float v = dot(r.d, qvec) * inv_det;
    if (v < 0.0f || u + v > 1.0f) return false;
    float t = dot(e2, qvec) * inv_det;
    if (t < r.tmin || t > r.tmax) return false;
    h.t = t;
    h.point = r.o + t * r.d;
    float w = 1.0f - u - v;
    v3 nv = cv3(tri.n0) * w + cv3(tri.n1) * u + cv3(tri.n2) * v;
    float len2 = dot(nv, nv);
    if (len2 > 1e-12f) h.normal = normalize(nv);
    else h.normal = normalize(cross(e1, e2));
    h.tangent = normalize(e1);
    h.uvx = tri.uv0.x * w + tri.uv1.x * u + tri.uv2.x * v;
    h.uvy = tri.uv0.y * w + tri.uv1.y * u + tri.uv2.y * v;
    if (dot(h.normal, r.d) > 0.0f) h.normal = -h.normal;
    return true;
}

// Cone frustum segment (reference primitives/cone.h:11-73): cone_solve holds
// the roots of a frustum once its frame is set up (the rest of hit_cone).
SHD bool cone_roots(const DRay& r, float A, float B, float C, v3 Dp, v3 Qp, v3 da, float u, float v, float L,
                    DHit& h) {
    float disc = B * B - 4.0f * A * C;
    if (disc < 0.0f) return false;
    float sq = sqrtf(disc);
    float inv2A = 1.0f / (2.0f * A);
    float s1 = (-B - sq) * inv2A;
    float s2 = (-B + sq) * inv2A;
    for (int i = 0; i < 2; i++) {
        float s = (i == 0) ? s1 : s2;
        if (s < r.tmin || s > r.tmax) continue;
        float tp = v + s * u;
        if (tp < 0.0f || tp > L) continue;
        h.t = s;
        h.point = r.o + s * r.d;
        h.normal = normalize(Qp + s * Dp);
        h.tangent = da;
        if (dot(h.normal, r.d) > 0.0f) h.normal = -h.normal;
        return true;
    }
    return false;
}
SHD bool cone_solve(const DRay& r, v3 Dp, v3 Qp, v3 da, float u, float v, float k, float r0, float L, DHit& h) {
    float A = dot(Dp, Dp) - k * k * u * u;
    float B = 2.0f * (dot(Dp, Qp) - k * r0 * u - k * k * v * u);
    float C = dot(Qp, Qp) - k * k * v * v - 2.0f * k * r0 * v - r0 * r0;
    return cone_roots(r, A, B, C, Dp, Qp, da, u, v, L, h);
}
// `smax`: a caller that only takes hits at distances up to smax (inclusive)
// passes it to skip a frustum whose axial coordinate tp = v + s*u, for s in
// [t_min, smax], stays outside [0, L]: in float, s*u and then v + (s*u)
// round monotonically, so tp over that range lies between its values at the
// two ends, and no root there can pass the axial test below (a NaN end
// never skips).  Hits the frustum could report beyond smax are not needed.
SHD bool hit_cone(const DRay& r, v3 p0, float r0, v3 p1, float r1, DHit& h, float smax) {
    v3 axis = p1 - p0;
    float L2 = dot(axis, axis);
    if (L2 < 1e-12f) return false;
    float L = sqrtf(L2);
    v3 da = axis / L;
    float k = (r1 - r0) / L;
    v3 Q = r.o - p0;
    float u = dot(r.d, da);
    float v = dot(Q, da);
    {
        const float ta = v + r.tmin * u, tb = v + smax * u;
        if ((ta < 0.0f && tb < 0.0f) || (ta > L && tb > L)) return false;
    }
    v3 Dp = r.d - u * da;
    v3 Qp = Q - v * da;
    return cone_solve(r, Dp, Qp, da, u, v, k, r0, L, h);
}
// hit_cone for the frustum from (0, y0, 0) to (0, y1, 0) (lathe profiles).
// The axis and its unit vector have exact zero x and z components (0 - 0,
// then 0 / L), so the products with them are left out: every value below
// equals hit_cone's up to the sign of a zero, which changes no comparison and
// no nonzero result, so the roots and the accepted hit are the same.
SHD bool hit_cone_y_div(const DRay& r, float y0, float r0, float y1, float r1, DHit& h, float smax) {
    const float ay = y1 - y0;
    const float L2 = ay * ay;
    if (L2 < 1e-12f) return false;
    const float L = sqrtf(L2);
    const float day = ay / L;
    const float k = (r1 - r0) / L;
    const float Qy = r.o.y - y0;
    const float u = r.d.y * day, v = Qy * day;
    {
        const float ta = v + r.tmin * u, tb = v + smax * u;
        if ((ta < 0.0f && tb < 0.0f) || (ta > L && tb > L)) return false;
    }
    const v3 Dp = mk3(r.d.x, r.d.y - u * day, r.d.z);
    const v3 Qp = mk3(r.o.x, Qy - v * day, r.o.z);
    return cone_solve(r, Dp, Qp, mk3(0.0f, day, 0.0f), u, v, k, r0, L, h);
}
// The ray's own terms of every y-axis frustum (ConeYRay, set once per lathe
// test): Dp = (d.x, 0, d.z) and Qp = (o.x, 0, o.z) whatever the frustum,
// since day = +-1 (below) makes d.y - u*day and Qy - v*day exact zeros (+0:
// x - x), so dot(Dp, Dp), dot(Dp, Qp) and dot(Qp, Qp) are the same sums of
// the same products for all of them.
struct ConeYRay {
    float DD, DQ, QQ;
};
SHD ConeYRay cone_y_ray(const DRay& r) {
    const v3 Dp = mk3(r.d.x, 0.0f, r.d.z), Qp = mk3(r.o.x, 0.0f, r.o.z);
    return {dot(Dp, Dp), dot(Dp, Qp), dot(Qp, Qp)};
}
// hit_cone_y_div with the square root and divisions of its set-up made
// exact: in binary floating point the square root of a rounded square is the
// magnitude (sqrtf(ay*ay) == |ay| whenever ay*ay is a normal float; checked
// over every float of [1, 4), which covers every exponent by scaling), so
// L = |ay| and day = ay / |ay| = +-1 exactly.  A square that overflows takes
// hit_cone_y_div.  Same roots, same accepted hit.
SHD bool hit_cone_y(const DRay& r, const ConeYRay& cr, float y0, float r0, float y1, float r1, DHit& h, float smax) {
    const float ay = y1 - y0;
    const float L2 = ay * ay;
    if (L2 < 1e-12f) return false;
    if (!(L2 <= 0x1.fffffep127f)) return hit_cone_y_div(r, y0, r0, y1, r1, h, smax);
    const float L = fabsf(ay);
    const float day = ay < 0.0f ? -1.0f : 1.0f;
    const float Qy = r.o.y - y0;
    const float u = r.d.y * day, v = Qy * day;
    {
        const float ta = v + r.tmin * u, tb = v + smax * u;
        if ((ta < 0.0f && tb < 0.0f) || (ta > L && tb > L)) return false;
    }
    const float k = (r1 - r0) / L;
    const v3 Dp = mk3(r.d.x, r.d.y - u * day, r.d.z);
    const v3 Qp = mk3(r.o.x, Qy - v * day, r.o.z);
    const float A = cr.DD - k * k * u * u;
    const float B = 2.0f * (cr.DQ - k * r0 * u - k * k * v * u);
    const float C = cr.QQ - k * k * v * v - 2.0f * k * r0 * v - r0 * r0;
    return cone_roots(r, A, B, C, Dp, Qp, mk3(0.0f, day, 0.0f), u, v, L, h);
}
// Per-frustum terms of a lathe profile (LatheFrustum, built once per scene
// by lathe_frustum): hit_cone_y's set-up, and the products of k and r0 that
// cone_solve forms the same way for every ray: k*k, k*r0, (2*k)*r0, r0*r0.
// Mode 1: the frustum is never hit (both radii <= 0, or L2 < 1e-12); mode 2:
// the square ay*ay overflows, and the frustum takes hit_cone_y_div.
struct LatheFrustum {
    float y0, L, day, k2;
    float kr0, tkr0, r02;
    int mode;
};
static_assert(sizeof(LatheFrustum) == 32, "one 32-byte record per frustum");
SHHD LatheFrustum lathe_frustum(float y0, float r0, float y1, float r1) {
    LatheFrustum f{};
    f.y0 = y0;
    const float ay = y1 - y0;
    const float L2 = ay * ay;
    if ((r0 <= 0.0f && r1 <= 0.0f) || L2 < 1e-12f) { f.mode = 1; return f; }
    if (!(L2 <= 0x1.fffffep127f)) { f.mode = 2; return f; }
    f.L = fabsf(ay);
    f.day = ay < 0.0f ? -1.0f : 1.0f;
    const float k = (r1 - r0) / f.L;
    f.k2 = k * k;
    f.kr0 = k * r0;
    f.tkr0 = 2.0f * k * r0;
    f.r02 = r0 * r0;
    return f;
}
// hit_cone_y from a frustum record of mode 0: the same operations on the
// same operands, so the same roots and the same accepted hit.
SHD bool hit_cone_y_rec(const DRay& r, const ConeYRay& cr, const LatheFrustum& f, DHit& h, float smax) {
    const float L = f.L, day = f.day;
    const float Qy = r.o.y - f.y0;
    const float u = r.d.y * day, v = Qy * day;
    {
        const float ta = v + r.tmin * u, tb = v + smax * u;
        if ((ta < 0.0f && tb < 0.0f) || (ta > L && tb > L)) return false;
    }
    const v3 Dp = mk3(r.d.x, r.d.y - u * day, r.d.z);
    const v3 Qp = mk3(r.o.x, Qy - v * day, r.o.z);
    const float A = cr.DD - f.k2 * u * u;
    const float B = 2.0f * (cr.DQ - f.kr0 * u - f.k2 * v * u);
    const float C = cr.QQ - f.k2 * v * v - f.tkr0 * v - f.r02;
    return cone_roots(r, A, B, C, Dp, Qp, mk3(0.0f, day, 0.0f), u, v, L, h);
}
// Half sphere on the side of `axis` (reference primitives/cone.h:78-121);
// leaves h.tangent untouched, as the reference does.
SHD bool hit_cap(const DRay& r, v3 c, float radius, v3 axis, DHit& h) {
    v3 oc = r.o - c;
    float a = dot(r.d, r.d);
    float b = 2.0f * dot(oc, r.d);
    float cc = dot(oc, oc) - radius * radius;
    float disc = b * b - 4.0f * a * cc;
    if (disc < 0.0f) return false;
    float sq = sqrtf(disc);
    float inv2a = 1.0f / (2.0f * a);
    float t1 = (-b - sq) * inv2a;
    float t2 = (-b + sq) * inv2a;
    for (int i = 0; i < 2; i++) {
        float t = (i == 0) ? t1 : t2;
        if (t < r.tmin || t > r.tmax) continue;
        v3 p = r.o + t * r.d;
        v3 n = normalize(p - c);
        if (dot(n, axis) < 0.0f) continue;
        h.t = t;
        h.point = p;
        h.normal = n;
        if (dot(h.normal, r.d) > 0.0f) h.normal = -h.normal;
        return true;
    }
    return false;
}
// Sphere-swept polyline (reference primitives/sweep.h:13-92).  `smax` as for
// hit_lathe: the frusta past the running best are skipped.  That is exact for
// the distance only: a cap leaves the tangent of the scratch record alone, so
// a cap hit carries the tangent of the last frustum that reported a hit in
// the same call, even one past the running best.  A caller that needs the
// tangent (the shading rebuild) passes skip = false and every frustum is
// solved.
SHD bool hit_strand(const DRay& r, const SweepStrand& st, const SweepPoint* pts, DHit& h, float smax,
                    bool skip = true) {
    if (st.point_count < 2) return false;
    bool any = false;
    float closest = r.tmax;
    DHit tmp;
    tmp.tangent = mk3(0, 0, 0);
    const SweepPoint* p = pts + st.point_offset;
    int count = st.point_count;
#define SHEEN_TAKE()                                                                     \
    if (tmp.t < closest) {                                                               \
        h = tmp; h.material_id = st.material_id; h.uvx = 0.0f; h.uvy = 0.0f;             \
        closest = tmp.t; any = true;                                                     \
    }
    if (hit_cap(r, cv3(p[0].position), p[0].radius, normalize(cv3(p[0].position) - cv3(p[1].position)), tmp)) { SHEEN_TAKE() }
    for (int i = 0; i < count - 1; i++) {
        if (hit_cone(r, cv3(p[i].position), p[i].radius, cv3(p[i + 1].position), p[i + 1].radius, tmp,
                     skip ? fminf(smax, closest) : r.tmax)) { SHEEN_TAKE() }
        if (i < count - 2) {
            if (hit_cap(r, cv3(p[i + 1].position), p[i + 1].radius,
                        normalize(cv3(p[i + 2].position) - cv3(p[i + 1].position)), tmp)) { SHEEN_TAKE() }
        }
    }
    if (hit_cap(r, cv3(p[count - 1].position), p[count - 1].radius,
                normalize(cv3(p[count - 1].position) - cv3(p[count - 2].position)), tmp)) { SHEEN_TAKE() }
#undef SHEEN_TAKE
    return any;
}
// Optical depth of one strand (reference primitives/sweep.h:97-148).
SHD float strand_depth(const DRay& r, const SweepStrand& st, const SweepPoint* pts, float sigma) {
    float tau = 0.0f;
    for (int i = 0; i < st.point_count - 1; i++) {
        const SweepPoint& p0 = pts[st.point_offset + i];
        const SweepPoint& p1 = pts[st.point_offset + i + 1];
        v3 u = cv3(p1.position) - cv3(p0.position);
        float A = dot(u, u);
        float add = 0.0f;
        if (!(A < 1e-12f)) {
            v3 v = r.d;
            v3 w = r.o - cv3(p0.position);
            float B = dot(u, v), C = dot(v, v), D = dot(u, w), E = dot(v, w);
            float denom = A * C - B * B;
            float s = (denom > 1e-9f) ? (B * E - C * D) / denom : 0.0f;
            s = fminf(fmaxf(s, 0.0f), 1.0f);
            float t = (B * s + E) / C;
            if (t < 0.0f) {
                t = 0.0f;
                s = fminf(fmaxf(-D / A, 0.0f), 1.0f);
            }
            if (!(t > r.tmax)) {
                v3 cl = w + u * s - v * t;
                float d2 = dot(cl, cl);
                float rad = p0.radius + (p1.radius - p0.radius) * s;
                float r2 = rad * rad;
                if (d2 < r2) add = sigma * (2.0f * sqrtf(fmaxf(r2 - d2, 0.0f)));
            }
        }
        tau += add;
        if (tau > 6.0f) break;
    }
    return tau;
}

// Superellipsoid (reference primitives/superellipsoid.h:12-161).
SHD float se_field(const Superellipsoid& s, v3 p) {
    float qx = (p.x - s.center.x) / s.a;
    float qy = (p.y - s.center.y) / s.b;
    float qz = (p.z - s.center.z) / s.c;
    return powf(fabsf(qx), s.n) + powf(fabsf(qy), s.n) + powf(fabsf(qz), s.n);
}
SHD v3 se_gradient(const Superellipsoid& s, v3 p) {
    float qx = (p.x - s.center.x) / s.a;
    float qy = (p.y - s.center.y) / s.b;
    float qz = (p.z - s.center.z) / s.c;
    float gx = s.n * powf(fabsf(qx), s.n - 1.0f) * (qx >= 0.0f ? 1.0f : -1.0f) / s.a;
    float gy = s.n * powf(fabsf(qy), s.n - 1.0f) * (qy >= 0.0f ? 1.0f : -1.0f) / s.b;
    float gz = s.n * powf(fabsf(qz), s.n - 1.0f) * (qz >= 0.0f ? 1.0f : -1.0f) / s.c;
    return {gx, gy, gz};
}
SHD void se_slab(const DRay& r, const Superellipsoid& se, float& tA, float& tB) {
    float bmnx = se.center.x - se.a, bmny = se.center.y - se.b, bmnz = se.center.z - se.c;
    float bmxx = se.center.x + se.a, bmxy = se.center.y + se.b, bmxz = se.center.z + se.c;
    float ix = 1.0f / r.d.x;
    float tx0 = (bmnx - r.o.x) * ix, tx1 = (bmxx - r.o.x) * ix;
    if (tx0 > tx1) { float t = tx0; tx0 = tx1; tx1 = t; }
    float iy = 1.0f / r.d.y;
    float ty0 = (bmny - r.o.y) * iy, ty1 = (bmxy - r.o.y) * iy;
    if (ty0 > ty1) { float t = ty0; ty0 = ty1; ty1 = t; }
    float iz = 1.0f / r.d.z;
    float tz0 = (bmnz - r.o.z) * iz, tz1 = (bmxz - r.o.z) * iz;
    if (tz0 > tz1) { float t = tz0; tz0 = tz1; tz1 = t; }
    tA = fmaxf(fmaxf(tx0, ty0), tz0);
    tB = fminf(fminf(tx1, ty1), tz1);
}
SHD bool se_span(const DRay& r, const Superellipsoid& se, float t0, float t1, float& te, float& tx) {
    {
        float tA, tB;
        se_slab(r, se, tA, tB);
        if (tB < t0 || tA > t1 || tB < tA) return false;
        t0 = fmaxf(t0, tA);
        t1 = fminf(t1, tB);
    }
    const float STEP = (t1 - t0) / 32.0f;
    float t_prev = t0;
    bool inside = se_field(se, r.o + t_prev * r.d) <= 1.0f;
    te = -1e30f; tx = -1e30f;
    bool fe = false, fx = false;
    if (inside) { te = t0; fe = true; }
    for (int i = 1; i <= 32 && !fx; i++) {
        float t = fminf(t0 + STEP * (float)i, t1);
        float F = se_field(se, r.o + t * r.d);
        if (inside && F > 1.0f) {
            float lo = t_prev, hi = t;
            for (int b = 0; b < 20; b++) {
                float mid = 0.5f * (lo + hi);
                if (se_field(se, r.o + mid * r.d) > 1.0f) hi = mid; else lo = mid;
            }
            tx = 0.5f * (lo + hi);
            if (tx > t1) tx = t1;
            fx = true;
            break;
        }
        t_prev = t;
        if (!inside && F <= 1.0f) {
            float lo = t_prev, hi = t;
            for (int b = 0; b < 20; b++) {
                float mid = 0.5f * (lo + hi);
                if (se_field(se, r.o + mid * r.d) <= 1.0f) hi = mid; else lo = mid;
            }
            te = 0.5f * (lo + hi);
            if (te < t0) te = t0;
            fe = true;
            inside = true;
        }
    }
    if (!fe) return false;
    if (!fx) tx = t1;
    return te <= tx;
}
// The entry distance se_span returns, found without the exit search: the
// entry is the first inside sample (the entry bisection starts with lo = hi
// and returns that sample), and te <= tx always holds (the exit lies at or
// after the last sample before it), so se_span succeeds exactly when an entry
// is found.  Samples are non-decreasing in distance, so once one lies beyond
// `smax` (a traversal's running best, inclusive) no later entry can be taken.
SHD bool se_entry(const DRay& r, const Superellipsoid& se, float t0, float t1, float smax, float& te) {
    {
        float tA, tB;
        se_slab(r, se, tA, tB);
        if (tB < t0 || tA > t1 || tB < tA) return false;
        t0 = fmaxf(t0, tA);
        t1 = fminf(t1, tB);
    }
    const float STEP = (t1 - t0) / 32.0f;
    if (se_field(se, r.o + t0 * r.d) <= 1.0f) { te = t0; return true; }
    for (int i = 1; i <= 32; i++) {
        const float t = fminf(t0 + STEP * (float)i, t1);
        if (t > smax) return false;
        if (se_field(se, r.o + t * r.d) <= 1.0f) { te = t; return true; }
    }
    return false;
}
// Attributes of a superellipsoid hit at distance te.
SHD void se_attrs(const DRay& r, const Superellipsoid& se, float te, DHit& h) {
    v3 P = r.o + te * r.d;
    h.t = te;
    h.point = P;
    h.normal = normalize(se_gradient(se, P));
    h.tangent = mk3(1, 0, 0);
    h.uvx = P.x; h.uvy = P.y;
    if (dot(h.normal, r.d) > 0.0f) h.normal = -h.normal;
}
// `attrs`: false when only the distance is needed (traversal).
SHD bool hit_se(const DRay& r, const Superellipsoid& se, DHit& h, float smax, bool attrs) {
    float t0, t1;
    se_slab(r, se, t0, t1);
    if (t1 < r.tmin || t0 > r.tmax || t1 < t0) return false;
    t0 = fmaxf(t0, r.tmin);
    t1 = fminf(t1, r.tmax);
    float te;
    if (!se_entry(r, se, t0, t1, smax, te)) return false;
    if (attrs) se_attrs(r, se, te, h);
    else h.t = te;
    return true;
}

// Height field (reference primitives/heightfield.h:13-183).
SHD float hf_hash(int xi, int zi, int seed) {
    uint32_t hh = uint32_t(xi) * 374761393u + uint32_t(zi) * 668265263u + uint32_t(seed) * 974634097u;
    hh = (hh ^ (hh >> 13)) * 1274126177u;
    hh = hh ^ (hh >> 16);
    return float(hh & 0xffffffu) / 16777216.0f;
}
SHD float hf_noise(float fx, float fz, int seed) {
    int x0 = (int)floorf(fx), z0 = (int)floorf(fz);
    float tx = fx - (float)x0, tz = fz - (float)z0;
    tx = tx * tx * (3.0f - 2.0f * tx);
    tz = tz * tz * (3.0f - 2.0f * tz);
    float v00 = hf_hash(x0, z0, seed), v10 = hf_hash(x0 + 1, z0, seed);
    float v01 = hf_hash(x0, z0 + 1, seed), v11 = hf_hash(x0 + 1, z0 + 1, seed);
    float a = v00 + (v10 - v00) * tx;
    float b = v01 + (v11 - v01) * tx;
    return a + (b - a) * tz;
}
SHD float hf_height(const HeightField& hf, float x, float z) {
    float lx = (x - hf.center.x) / hf.size_x + 0.5f;
    float lz = (z - hf.center.z) / hf.size_z + 0.5f;
    if (lx < 0.0f || lx > 1.0f || lz < 0.0f || lz > 1.0f) return 0.0f;
    if (hf.image) {
        if (hf.img_w < 1 || hf.img_h < 1) return 0.0f;
        float fu = lx * (float)(hf.img_w - 1);
        float fv = lz * (float)(hf.img_h - 1);
        int u0 = (int)fu, v0 = (int)fv;
        int u1 = (u0 < hf.img_w - 1) ? u0 + 1 : u0;
        int v1 = (v0 < hf.img_h - 1) ? v0 + 1 : v0;
        float tx = fu - (float)u0, ty = fv - (float)v0;
        float h00 = (float)hf.image[v0 * hf.img_w + u0] / 255.0f, h10 = (float)hf.image[v0 * hf.img_w + u1] / 255.0f;
        float h01 = (float)hf.image[v1 * hf.img_w + u0] / 255.0f, h11 = (float)hf.image[v1 * hf.img_w + u1] / 255.0f;
        float hv = h00 * (1.0f - tx) * (1.0f - ty) + h10 * tx * (1.0f - ty) + h01 * (1.0f - tx) * ty + h11 * tx * ty;
        return hv * hf.height;
    }
    float fx = lx * hf.freq, fz = lz * hf.freq;
    float amp = 0.5f, sum = 0.0f, norm = 0.0f;
    for (int o = 0; o < hf.octaves; o++) {
        sum += amp * hf_noise(fx, fz, hf.seed + o * 131);
        norm += amp;
        amp *= 0.5f;
        fx *= 2.0f; fz *= 2.0f;
    }
    float t = sum / norm;
    float detail = hf_noise(lx * 7.0f, lz * 7.0f, hf.seed + 7777) * 0.06f;
    t = fminf(t + detail - 0.03f, 1.0f);
    return t * hf.height;
}
// Attributes of a height-field hit at distance tt, point P = r.o + tt * r.d.
SHD void hf_attrs(const DRay& r, const HeightField& hf, float tt, v3 P, DHit& h) {
    h.t = tt;
    h.point = P;
    float e = fmaxf(hf.size_x, hf.size_z) / 256.0f;
    float hxp = hf_height(hf, P.x + e, P.z);
    float hxm = hf_height(hf, P.x - e, P.z);
    float hzp = hf_height(hf, P.x, P.z + e);
    float hzm = hf_height(hf, P.x, P.z - e);
    h.normal = normalize(mk3(-(hxp - hxm) / (2.0f * e), 1.0f, -(hzp - hzm) / (2.0f * e)));
    h.tangent = mk3(1, 0, 0);
    h.uvx = P.x; h.uvy = P.z;
    if (dot(h.normal, r.d) > 0.0f) h.normal = -h.normal;
}
// `smax`: only hits up to smax (inclusive) are needed: a crossing found at a
// march step brackets the hit in [max(t - step, t0), t] (the bisection keeps
// it there and the midpoint of values at or above a bound is at or above
// it), and the steps only increase, so the march stops once that lower end
// passes smax.  `attrs`: false when only the distance is needed (traversal).
SHD bool hit_hf(const DRay& r, const HeightField& hf, DHit& h, float smax, bool attrs) {
    float bmnx = hf.center.x - 0.5f * hf.size_x, bmny = hf.center.y, bmnz = hf.center.z - 0.5f * hf.size_z;
    float bmxx = hf.center.x + 0.5f * hf.size_x, bmxy = hf.center.y + hf.height, bmxz = hf.center.z + 0.5f * hf.size_z;
    float t0, t1;
    {
        float ix = 1.0f / r.d.x, iy = 1.0f / r.d.y, iz = 1.0f / r.d.z;
        float tx0 = (bmnx - r.o.x) * ix, tx1 = (bmxx - r.o.x) * ix;
        float ty0 = (bmny - r.o.y) * iy, ty1 = (bmxy - r.o.y) * iy;
        float tz0 = (bmnz - r.o.z) * iz, tz1 = (bmxz - r.o.z) * iz;
        if (tx0 > tx1) { float t = tx0; tx0 = tx1; tx1 = t; }
        if (ty0 > ty1) { float t = ty0; ty0 = ty1; ty1 = t; }
        if (tz0 > tz1) { float t = tz0; tz0 = tz1; tz1 = t; }
        t0 = fmaxf(fmaxf(tx0, ty0), tz0);
        t1 = fminf(fminf(tx1, ty1), tz1);
        if (t1 < r.tmin || t0 > r.tmax || t1 < t0) return false;
        t0 = fmaxf(t0, r.tmin);
        t1 = fminf(t1, r.tmax);
    }
    float diag = sqrtf(hf.size_x * hf.size_x + hf.size_z * hf.size_z);
    float step = fmaxf(diag / 64.0f, 1e-4f);
    float t = t0;
    bool inside = false;
    for (int i = 0; i < 256; i++) {
        {
            float lo0 = t - step;
            if (lo0 < t0) lo0 = t0;
            if (lo0 > smax) break;
        }
        v3 p = r.o + t * r.d;
        bool iy = (p.y <= hf_height(hf, p.x, p.z) + hf.center.y);
        if (iy != inside) {
            float lo = t - step, hi = t;
            if (lo < t0) lo = t0;
            for (int b = 0; b < 10; b++) {
                float mid = 0.5f * (lo + hi);
                v3 pm = r.o + mid * r.d;
                if ((pm.y <= hf_height(hf, pm.x, pm.z) + hf.center.y) == iy) hi = mid; else lo = mid;
            }
            float tt = 0.5f * (lo + hi);
            v3 P = r.o + tt * r.d;
            if (tt >= r.tmin && tt <= r.tmax) {
                if (attrs) hf_attrs(r, hf, tt, P, h);
                else h.t = tt;
                return true;
            }
            break;
        }
        inside = iy;
        t += step;
        if (t > t1) break;
    }
    return false;
}

// Lathe: cone segments on the Y axis (reference primitives/lathe.h:14-55).
// `smax`: only hits at distances up to smax are needed (a traversal's
// running best; r.tmax for the full result); the closest hit at or below it
// is the same as without the bound.
SHD bool hit_lathe(const DRay& r, const Lathe& l, const LathePoint* pts, DHit& h, float smax) {
    if (l.point_count < 2) return false;
    bool any = false;
    float closest = r.tmax;
    DHit tmp;
    const ConeYRay cr = cone_y_ray(r);
    for (int i = 0; i < l.point_count - 1; i++) {
        const LathePoint& p0 = pts[l.point_offset + i];
        const LathePoint& p1 = pts[l.point_offset + i + 1];
        if (p0.radius <= 0.0f && p1.radius <= 0.0f) continue;
        if (hit_cone_y(r, cr, p0.y, p0.radius, p1.y, p1.radius, tmp, fminf(smax, closest)) && tmp.t < closest) {
            tmp.point = r.o + tmp.t * r.d;
            h = tmp;
            h.material_id = l.material_id;
            h.uvx = 0.0f; h.uvy = 0.0f;
            closest = tmp.t;
            any = true;
        }
    }
    return any;
}

// Prism (reference primitives/prism.h:18-134).
SHD bool prism_contour(float x, float z, const PrismPoint* pts, int base, int count) {
    bool inside = false;
    int j = count - 1;
    for (int i = 0; i < count; i++) {
        const PrismPoint& pi = pts[base + i];
        const PrismPoint& pj = pts[base + j];
        if ((pi.z > z) != (pj.z > z)) {
            float xint = (pj.x - pi.x) * (z - pi.z) / (pj.z - pi.z) + pi.x;
            if (x < xint) inside = !inside;
        }
        j = i;
    }
    return inside;
}
SHD bool prism_inside(float x, float z, const Prism& p, const PrismPoint* all) {
    int base = p.point_offset, cnt = p.point_count, i = 0;
    bool inside = false;
    while (i < cnt) {
        int start = i, j = i + 1;
        while (j < cnt) {
            if (all[base + j].x == all[base + start].x && all[base + j].z == all[base + start].z) break;
            j++;
        }
        if (j - start >= 3 && prism_contour(x, z, all, base + start, j - start)) inside = !inside;
        i = (j < cnt) ? j + 1 : j;
    }
    return inside;
}
SHD bool hit_prism(const DRay& r, const Prism& p, const PrismPoint* pts, DHit& h) {
    float y0 = p.center.y + p.y_min;
    float y1 = p.center.y + p.y_max;
    float ty0 = (y0 - r.o.y) / r.d.y;
    float ty1 = (y1 - r.o.y) / r.d.y;
    if (ty0 > ty1) { float t = ty0; ty0 = ty1; ty1 = t; }
    if (fmaxf(ty0, r.tmin) > fminf(ty1, r.tmax)) return false;
    bool any = false;
    float closest = r.tmax;
    for (int face = 0; face < 2; face++) {
        float yf = (face == 0) ? y0 : y1;
        float t = (yf - r.o.y) / r.d.y;
        if (t < r.tmin || t > r.tmax) continue;
        v3 P = r.o + t * r.d;
        if (!prism_inside(P.x - p.center.x, P.z - p.center.z, p, pts)) continue;
        if (t < closest) {
            h.t = t; h.point = P;
            h.normal = mk3(0.0f, (face == 0) ? -1.0f : 1.0f, 0.0f);
            if (dot(h.normal, r.d) > 0.0f) h.normal = -h.normal;
            h.tangent = mk3(1, 0, 0);
            h.material_id = p.material_id; h.uvx = 0.0f; h.uvy = 0.0f;
            closest = t; any = true;
        }
    }
    float dx = r.d.x, dy = r.d.y, dz = r.d.z, ox = r.o.x, oz = r.o.z;
    int base = p.point_offset, cnt = p.point_count;
    for (int i = 0; i < cnt; i++) {
        int j = (i + 1) % cnt;
        const PrismPoint& pi = pts[base + i];
        const PrismPoint& pj = pts[base + j];
        if (pi.x == pj.x && pi.z == pj.z) continue;
        float ax = pi.x + p.center.x, az = pi.z + p.center.z;
        float bx = pj.x + p.center.x, bz = pj.z + p.center.z;
        float ex = bx - ax, ez = bz - az;
        float denom = ex * dz - ez * dx;
        if (fabsf(denom) < 1e-9f) continue;
        float s = (dx * (az - oz) - dz * (ax - ox)) / denom;
        if (s < -0.001f || s > 1.001f) continue;
        float t;
        if (fabsf(dx) > 1e-9f) t = (ax + s * ex - ox) / dx;
        else if (fabsf(dz) > 1e-9f) t = (az + s * ez - oz) / dz;
        else continue;
        if (t < r.tmin || t > r.tmax) continue;
        float yy = r.o.y + dy * t;
        if (yy < y0 - 0.001f || yy > y1 + 0.001f) continue;
        if (t < closest) {
            h.t = t; h.point = r.o + t * r.d;
            h.normal = normalize(mk3(ez, 0.0f, -ex));
            h.tangent = mk3(0, 1, 0);
            if (dot(h.normal, r.d) > 0.0f) h.normal = -h.normal;
            h.material_id = p.material_id; h.uvx = 0.0f; h.uvy = 0.0f;
            closest = t; any = true;
        }
    }
    return any;
}

// CSG interval algebra (reference integrator/csg.h:22-186), evaluated
// post-order with an explicit stack instead of recursion.
struct CSpan { float te, tx; int mat, leaf, kind; };
#define CSG_MAX_DEPTH 16

SHD void csg_combine(const CSpan* A, int nA, const CSpan* B, int nB, int op, CSpan* out, int& nOut) {
    nOut = 0;
    if (op == 0) {
        CSpan tmp[16];
        int n = 0;
        for (int i = 0; i < nA && n < 16; i++) tmp[n++] = A[i];
        for (int i = 0; i < nB && n < 16; i++) tmp[n++] = B[i];
        for (int i = 1; i < n; i++) {
            CSpan key = tmp[i];
            int j = i - 1;
            while (j >= 0 && tmp[j].te > key.te) { tmp[j + 1] = tmp[j]; j--; }
            tmp[j + 1] = key;
        }
        for (int i = 0; i < n && nOut < 8; i++) {
            if (nOut == 0 || tmp[i].te > out[nOut - 1].tx) out[nOut++] = tmp[i];
            else if (tmp[i].tx > out[nOut - 1].tx) out[nOut - 1].tx = tmp[i].tx;
        }
    } else if (op == 1) {
        for (int i = 0; i < nA; i++)
            for (int j = 0; j < nB; j++) {
                float t0 = fmaxf(A[i].te, B[j].te), t1 = fminf(A[i].tx, B[j].tx);
                if (t0 < t1 && nOut < 8) { out[nOut] = A[i]; out[nOut].te = t0; out[nOut].tx = t1; nOut++; }
            }
    } else if (op == 2) {
        for (int i = 0; i < nA; i++) {
            float cur = A[i].te;
            for (int j = 0; j < nB; j++) {
                float b0 = B[j].te, b1 = B[j].tx;
                if (b1 <= cur) continue;
                if (b0 >= A[i].tx) break;
                if (b0 > cur && nOut < 8) { out[nOut] = A[i]; out[nOut].te = cur; out[nOut].tx = b0; nOut++; }
                cur = fmaxf(cur, b1);
                if (cur >= A[i].tx) break;
            }
            if (cur < A[i].tx && nOut < 8) { out[nOut] = A[i]; out[nOut].te = cur; out[nOut].tx = A[i].tx; nOut++; }
        }
    }
}
