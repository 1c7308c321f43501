// Constraint math shared by host and device code.  Every function is a fixed
// sequence of IEEE single operations (the library is built with
// -ffp-contract=off), so host and device results agree bit-for-bit.
#pragma once

#include <hip/hip_runtime.h>

namespace pbd {

#define PBD_HD __host__ __device__ __forceinline__

PBD_HD float atan2_hd(float y, float x) {
    float ax = fabsf(x), ay = fabsf(y);
    float mx = ax > ay ? ax : ay;
    float mn = ax > ay ? ay : ax;
    if (mx == 0.0f) return 0.0f;
    float a = mn / mx;
    float s = a * a;
    float r = -0x1.09b858p-8f;
    r = r * s + 0x1.6633e4p-6f;
    r = r * s + -0x1.ca08a6p-5f;
    r = r * s + 0x1.8af1c4p-4f;
    r = r * s + -0x1.1cd946p-3f;
    r = r * s + 0x1.988174p-3f;
    r = r * s + -0x1.554c3ap-2f;
    r = r * s + 0x1.ffffeap-1f;
    r = r * a;
    if (ay > ax) r = 1.57079637f - r;
    if (x < 0.0f) r = 3.14159274f - r;
    if (y < 0.0f) r = -r;
    return r;
}

struct V3 {
    float x, y, z;
};
PBD_HD V3 mk(float x, float y, float z) { return V3{x, y, z}; }
PBD_HD V3 sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
PBD_HD float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
PBD_HD V3 cross(V3 a, V3 b) {
    return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
PBD_HD V3 scale(float s, V3 a) { return V3{s * a.x, s * a.y, s * a.z}; }

// Correctly rounded square root for x in [2^-96, 2^96]: the hardware
// estimate s (within 1 ulp) is replaced by its neighbour s - ulp or s + ulp
// when the fused residual x - s' * s shows that neighbour is nearer
// (Markstein's correction).  In that range it is the sequence a correctly
// rounded sqrtf runs without its rescaling of small inputs, so the result
// equals sqrtf bit for bit; an exhaustive test over every float of the range
// checks it.
__device__ __forceinline__ float sqrt_rn(float x) {
    float s = __builtin_amdgcn_sqrtf(x);
    const float sd = __int_as_float(__float_as_int(s) - 1);
    const float su = __int_as_float(__float_as_int(s) + 1);
    const float rd = __builtin_fmaf(-sd, s, x);
    const float ru = __builtin_fmaf(-su, s, x);
    s = rd <= 0.0f ? sd : s;
    s = ru > 0.0f ? su : s;
    return s;
}

// Correctly rounded 1/d for d in [2^-48, 2^48]: one Newton step on the
// hardware reciprocal r (within 1 ulp), r + r (1 - d r) with the residual
// fused and exact (Markstein 1990).  Equal to the IEEE quotient 1.0f / d on
// every float of the range, checked exhaustively.
__device__ __forceinline__ float rcp_rn(float d) {
    const float r = __builtin_amdgcn_rcpf(d);
    const float e = __builtin_fmaf(-d, r, 1.0f);
    return __builtin_fmaf(e, r, r);
}

// Distance projection on two (x, y, z, w) records: the reference's CPU
// formula r = normalize(p0-p1), dp0 = -w0/(w0+w1)(d-L) r, dp1 = +w1/(w0+w1)(d-L) r
// (PBD_cuda.cu:1105-1124) scaled by k (PBD_cuda.cu:1282-1287).
__device__ __forceinline__ void project_distance(float4& a, float4& b, float rest, float k) {
    float w0 = a.w, w1 = b.w;
    float ws = w0 + w1;
    if (ws == 0.0f) return;
    float vx = a.x - b.x, vy = a.y - b.y, vz = a.z - b.z;
    const float d2 = vx * vx + vy * vy + vz * vz;
    float d, inv;
    if (d2 >= 0x1p-96f && d2 <= 0x1p96f) {
        // lengths in [2^-48, 2^48]: the short forms, equal to sqrtf and 1.0f / d there
        d = sqrt_rn(d2);
        inv = rcp_rn(d);
    } else {
        d = sqrtf(d2);
        if (d == 0.0f) return;
        inv = 1.0f / d;
    }
    float rx = vx * inv, ry = vy * inv, rz = vz * inv;
    float c = d - rest;
    // -w0/ws and w1/ws; equal or zero weights give exactly representable
    // quotients (-1/2, 1/2, -0, 1, -1, 0), which IEEE division returns
    // exactly, so the shortcut is bit-identical to dividing
    float q0, q1;
    if (w0 == w1) {
        q0 = -0.5f;
        q1 = 0.5f;
    } else if (w0 == 0.0f) {
        q0 = -0.0f;
        q1 = 1.0f;
    } else if (w1 == 0.0f) {
        q0 = -1.0f;
        q1 = 0.0f;
    } else {
        q0 = -w0 / ws;
        q1 = w1 / ws;
    }
    float s0 = q0 * c;
    float s1 = q1 * c;
    float ax = s0 * rx, ay = s0 * ry, az = s0 * rz;
    float bx = s1 * rx, by = s1 * ry, bz = s1 * rz;
    // (s * r) * k with k == 1 is (s * r) exactly: the full stiffness of the
    // configs skips six multiplies
    if (k != 1.0f) {
        ax = ax * k;
        ay = ay * k;
        az = az * k;
        bx = bx * k;
        by = by * k;
        bz = bz * k;
    }
    a.x = a.x + ax;
    a.y = a.y + ay;
    a.z = a.z + az;
    b.x = b.x + bx;
    b.y = b.y + by;
    b.z = b.z + bz;
}

// Dihedral terms for hinge (e0,e1) with wings o0, o1 (Bridson et al. 2003):
// n1 = (o0-e0)x(o0-e1), n2 = (o1-e1)x(o1-e0), E = e1-e0.
struct Dih {
    V3 E, n1, n2;
    float n1sq, n2sq, elen;
    bool ok;
};

PBD_HD Dih dih_terms(V3 e0, V3 e1, V3 o0, V3 o1) {
    Dih d;
    d.E = sub(e1, e0);
    d.n1 = cross(sub(o0, e0), sub(o0, e1));
    d.n2 = cross(sub(o1, e1), sub(o1, e0));
    d.n1sq = dot(d.n1, d.n1);
    d.n2sq = dot(d.n2, d.n2);
    d.elen = sqrtf(dot(d.E, d.E));
    d.ok = (d.n1sq > 1e-24f) && (d.n2sq > 1e-24f) && (d.elen > 1e-12f);
    return d;
}

PBD_HD float dih_angle(const Dih& d) {
    V3 c = cross(d.n2, d.n1);
    float sn = dot(c, d.E) / d.elen;
    float cs = dot(d.n1, d.n2);
    return atan2_hd(sn, cs);
}

PBD_HD float dihedral_hd(V3 e0, V3 e1, V3 o0, V3 o1) {
    Dih d = dih_terms(e0, e1, o0, o1);
    if (!d.ok) return 0.0f;
    return dih_angle(d);
}

// Bending projection: C = theta - theta0, dp_i = -k w_i C / sum_j w_j |u_j|^2 u_i
// with u = grad theta (Mueller et al. 2007 generic step; gradient from
// Bridson et al. 2003).
PBD_HD void project_bending(float4& q0, float4& q1, float4& q2, float4& q3, float rest, float k) {
    V3 e0 = mk(q0.x, q0.y, q0.z), e1 = mk(q1.x, q1.y, q1.z);
    V3 o0 = mk(q2.x, q2.y, q2.z), o1 = mk(q3.x, q3.y, q3.z);
    Dih d = dih_terms(e0, e1, o0, o1);
    if (!d.ok) return;
    float th = dih_angle(d);
    float s1 = d.elen / d.n1sq;
    float s2 = d.elen / d.n2sq;
    float a1 = (dot(sub(o0, e1), d.E) / d.elen) / d.n1sq;
    float a2 = (dot(sub(o1, e1), d.E) / d.elen) / d.n2sq;
    float b1 = (dot(sub(o0, e0), d.E) / d.elen) / d.n1sq;
    float b2 = (dot(sub(o1, e0), d.E) / d.elen) / d.n2sq;
    V3 u0 = V3{a1 * d.n1.x + a2 * d.n2.x, a1 * d.n1.y + a2 * d.n2.y, a1 * d.n1.z + a2 * d.n2.z};
    V3 u1 = V3{-(b1 * d.n1.x) - b2 * d.n2.x, -(b1 * d.n1.y) - b2 * d.n2.y, -(b1 * d.n1.z) - b2 * d.n2.z};
    V3 u2 = scale(s1, d.n1);
    V3 u3 = scale(s2, d.n2);
    float den = 0.0f;
    den = den + q0.w * dot(u0, u0);
    den = den + q1.w * dot(u1, u1);
    den = den + q2.w * dot(u2, u2);
    den = den + q3.w * dot(u3, u3);
    if (!(den > 1e-20f)) return;
    float s = ((th - rest) / den) * k;
    float f;
    f = s * q0.w;
    q0.x = q0.x - f * u0.x; q0.y = q0.y - f * u0.y; q0.z = q0.z - f * u0.z;
    f = s * q1.w;
    q1.x = q1.x - f * u1.x; q1.y = q1.y - f * u1.y; q1.z = q1.z - f * u1.z;
    f = s * q2.w;
    q2.x = q2.x - f * u2.x; q2.y = q2.y - f * u2.y; q2.z = q2.z - f * u2.z;
    f = s * q3.w;
    q3.x = q3.x - f * u3.x; q3.y = q3.y - f * u3.y; q3.z = q3.z - f * u3.z;
}

// Analytic colliders.  Sphere: SolverPBD::ColliWithShpGrd / ColliderSphere /
// GenerateMoveVectorSphere (PBD_cuda.cu:1519-1573) with a parameterised centre
// and radius and the distance in sqrtf form; box (new): a point strictly
// inside is moved to the nearest face.  Pinned points (w == 0) are skipped.
PBD_HD void collide_point(float4& q, int sphere_on, const float* sc, float sr, int box_on,
                          const float* lo, const float* hi) {
    if (q.w == 0.0f) return;
    if (sphere_on) {
        float dx = q.x - sc[0], dy = q.y - sc[1], dz = q.z - sc[2];
        float d = sqrtf(dx * dx + dy * dy + dz * dz);
        if (!(d - sr > 0.001f) && d > 0.0f) {
            float md = sr - d;
            float ux = (q.x - sc[0]) / d;
            float uy = (q.y - sc[1]) / d;
            float uz = (q.z - sc[2]) / d;
            q.x = q.x + ux * md;
            q.y = q.y + uy * md;
            q.z = q.z + uz * md;
        }
    }
    if (box_on) {
        float c[3] = {q.x, q.y, q.z};
        bool inside = (c[0] > lo[0]) & (c[0] < hi[0]) & (c[1] > lo[1]) & (c[1] < hi[1]) &
                      (c[2] > lo[2]) & (c[2] < hi[2]);
        if (inside) {
            int bk = 0, bs = 0;
            float best = c[0] - lo[0];
            for (int k = 0; k < 3; k++) {
                float dl = c[k] - lo[k], dh = hi[k] - c[k];
                if (dl < best) { best = dl; bk = k; bs = 0; }
                if (dh < best) { best = dh; bk = k; bs = 1; }
            }
            c[bk] = bs ? hi[bk] : lo[bk];
            q.x = c[0];
            q.y = c[1];
            q.z = c[2];
        }
    }
}

}  // namespace pbd
