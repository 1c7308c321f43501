/*
 * pbd_oracle.c -- CPU oracle for the PBD cloth substep (TEST INFRASTRUCTURE).
 *
 * A plain-C restatement of the reference algorithms.  Reference paths are
 * relative to PositionBasedDynamicsCUDA/PositionBasedDynamicsCUDA/ ("PBD/").
 * Only tests/, __graft_entry__.smoke() and bench.py (cpu_baseline) use this
 * file; the product library never links it.
 *
 * Floating-point discipline: compiled with -ffp-contract=off and no fast-math,
 * so every expression below is a sequence of IEEE single/double operations in
 * the order written.  The HIP kernels evaluate the same sequences, which is
 * what makes bit-exact CPU/GPU comparison possible.
 *
 * Pinning summary (what checks each part independently of this file):
 *   grid/tris/pairs ......... closed-form counts and index formulas (tests)
 *   rest lengths ............ closed form (spacing, spacing*sqrt2) + powf/sqrtf agreement
 *   atan2 / dihedral ........ libm atan2 within 2e-7; FD gradient in FP64; rigid-motion invariance
 *   adjacency ............... brute-force set equality
 *   colouring ............... groundTruthTest KAT (PBD/PBD_cuda.cu:871-881), validity, first-fit
 *   colour sort / worksets .. stable-sort definition (numpy argsort kind=stable)
 *   advect / integrate ...... closed form free fall
 *   distance projection ..... two-vertex analytic solution
 *   bending projection ...... zero at rest, momentum conservation, FD gradient
 *   colliders ............... closed form
 *   residual ................ brute force in numpy (FP64)
 *   broadphase .............. brute force over all triangle pairs
 *   CCD cubic ............... numpy.roots in FP64; constructed trajectories with known TOI
 *   response ................ plane-distance invariant
 *   MLP / Adam .............. FD gradients; Adam closed form on a quadratic
 */
#include "pbd_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* small float3 helpers (semantics of PBD/helper_math.h:1248,1309,1420;     */
/* host normalize = v * (1/sqrtf(dot(v,v)))  PBD/helper_math.h:62-65)        */
/* ------------------------------------------------------------------------ */
static inline void v3sub(const float* a, const float* b, float* r) {
    r[0] = a[0] - b[0];
    r[1] = a[1] - b[1];
    r[2] = a[2] - b[2];
}
static inline float v3dot(const float* a, const float* b) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
static inline void v3cross(const float* a, const float* b, float* r) {
    float x = a[1] * b[2] - a[2] * b[1];
    float y = a[2] * b[0] - a[0] * b[2];
    float z = a[0] * b[1] - a[1] * b[0];
    r[0] = x;
    r[1] = y;
    r[2] = z;
}
static inline void v3normalize(const float* a, float* r) {
    float inv = 1.0f / sqrtf(v3dot(a, a));
    r[0] = a[0] * inv;
    r[1] = a[1] * inv;
    r[2] = a[2] * inv;
}

/* ------------------------------------------------------------------------ */
/* RNG: splitmix64 (Steele, Lea, Flood 2014).  uniform_pm1 maps the top 24   */
/* bits to a float in [-1, 1).                                               */
/* ------------------------------------------------------------------------ */
uint64_t orc_splitmix64(uint64_t* state) {
    uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

float orc_uniform_pm1(uint64_t* state) {
    uint64_t r = orc_splitmix64(state) >> 40;           /* 24 bits */
    float u = (float)r * (1.0f / 16777216.0f);          /* [0,1) exact */
    return 2.0f * u - 1.0f;
}

/* PBDObject::initPosition  PBD/PBD_cuda.cu:831-852 (cord = (0,0)) */
void orc_grid_positions(int rx, int ry, float sx, float sy, float* pos) {
    float li = sx / (rx - 1);
    float hi = sy / (ry - 1);
    int k = 0;
    for (int i = 0; i < ry; i++) {
        for (int j = 0; j < rx; j++) {
            pos[3 * k + 0] = 0.0f + j * li;
            pos[3 * k + 1] = 0.0f;
            pos[3 * k + 2] = 0.0f + i * hi;
            k++;
        }
    }
}

/* Seeded jitter: one uniform draw per coordinate in vertex order. */
void orc_jitter(float* pos, int nv, float amp, uint64_t seed) {
    uint64_t s = seed;
    for (int i = 0; i < 3 * nv; i++) pos[i] = pos[i] + amp * orc_uniform_pm1(&s);
}

/* PBDObject::initMeshTopolIndices  PBD/PBD_cuda.cu:854-869 */
int orc_grid_tris(int rx, int ry, int* tris) {
    int num = rx * ry, t = 0;
    for (int i = 0; i < num - rx; i++) {
        if (i % rx == rx - 1) continue;
        if (tris) {
            int* q = tris + 3 * t;
            q[0] = i; q[1] = i + rx; q[2] = i + rx + 1;
            q[3] = i; q[4] = i + rx + 1; q[5] = i + 1;
        }
        t += 2;
    }
    return t;
}

/* ConstraintPBD::InitDistanceIndices  PBD/PBD_cuda.cu:19-54 */
int orc_grid_distance_pairs(int rx, int ry, int* pairs) {
    int num = rx * ry, e = 0;
#define PUSH(a, b) do { if (pairs) { pairs[2 * e] = (a); pairs[2 * e + 1] = (b); } e++; } while (0)
    for (int i = 0; i < num - rx; i++) {
        if (i % rx == 0) {
            PUSH(i, i + 1);
            PUSH(i, i + rx);
        } else if (i % rx == rx - 1) {
            PUSH(i, i + rx);
            PUSH(i, i + rx - 1);
        } else {
            PUSH(i, i + 1);
            PUSH(i, i + rx);
            PUSH(i, i + rx - 1);
        }
    }
    for (int i = num - rx; i < num - 1; ++i) PUSH(i, i + 1);
#undef PUSH
    return e;
}

/* Distance  PBD/PBD_cuda.cu:3-6 (the reference's powf form; used for rest
 * lengths exactly as InitDistanceInfo does, PBD/PBD_cuda.cu:97). */
float orc_distance_powf(const float* a, const float* b) {
    return powf(powf((a[0] - b[0]), 2) + powf((a[1] - b[1]), 2) + powf((a[2] - b[2]), 2), 0.5);
}

/* Same quantity with multiplies and a correctly rounded sqrtf: the form the
 * solver uses on CPU and GPU (agreement with the powf form is tested). */
float orc_distance(const float* a, const float* b) {
    float dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
    return sqrtf(dx * dx + dy * dy + dz * dz);
}

/* ConstraintPBD::InitDistanceInfo  PBD/PBD_cuda.cu:56-108 (rest length) */
void orc_rest_lengths(const float* pos, const int* pairs, int ne, float* rest) {
    for (int i = 0; i < ne; i++)
        rest[i] = orc_distance_powf(pos + 3 * pairs[2 * i], pos + 3 * pairs[2 * i + 1]);
}

/* ------------------------------------------------------------------------ */
/* Bending.  The reference's InitBendingConstr is an empty stub             */
/* (PBD/PBD_cuda.cu:110-123).  New: one dihedral constraint per interior     */
/* triangle edge, angle and gradient after Bridson, Marino, Fedkiw 2003      */
/* ("Simulation of clothing with folds and wrinkles", sec. 4.1).             */
/* Quad layout: (e0, e1, o0, o1) = hinge edge as it appears in the first     */
/* triangle's winding, opposite vertex of the first triangle, opposite       */
/* vertex of the second.  Emitted in order of the second sighting while      */
/* scanning triangles in index order.                                        */
/* ------------------------------------------------------------------------ */
typedef struct { long long key; int a, b, o, used; } edge_slot;

static unsigned long long mix64(unsigned long long k) {
    k ^= k >> 33; k *= 0xff51afd7ed558ccdull; k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull; k ^= k >> 33;
    return k;
}

int orc_bending_quads(const int* tris, int nt, int* quads) {
    size_t cap = 1;
    while (cap < (size_t)(3 * nt) * 2 + 16) cap <<= 1;
    edge_slot* tab = (edge_slot*)calloc(cap, sizeof(edge_slot));
    int nb = 0;
    for (int t = 0; t < nt; t++) {
        for (int k = 0; k < 3; k++) {
            int a = tris[3 * t + k], b = tris[3 * t + (k + 1) % 3], o = tris[3 * t + (k + 2) % 3];
            int lo = a < b ? a : b, hi = a < b ? b : a;
            long long key = ((long long)lo << 32) | (unsigned)hi;
            size_t h = (size_t)mix64((unsigned long long)key) & (cap - 1);
            while (tab[h].used && tab[h].key != key) h = (h + 1) & (cap - 1);
            if (!tab[h].used) {
                tab[h].used = 1; tab[h].key = key; tab[h].a = a; tab[h].b = b; tab[h].o = o;
            } else if (tab[h].used == 1) {
                if (quads) {
                    quads[4 * nb + 0] = tab[h].a;
                    quads[4 * nb + 1] = tab[h].b;
                    quads[4 * nb + 2] = tab[h].o;
                    quads[4 * nb + 3] = o;
                }
                nb++;
                tab[h].used = 2; /* non-manifold third sightings are ignored */
            }
        }
    }
    free(tab);
    return nb;
}

/* atan2 from IEEE single operations only, so CPU and GPU agree bit-for-bit:
 * octant reduction to a = min/max in [0,1], then atan(a) = a*P(a^2) with an
 * 8-term polynomial fitted by Lawson-weighted least squares on [0,1]
 * (max abs error 1.4e-7 rad in single precision). */
float orc_atan2f(float y, float x) {
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

/* Shared geometric terms of the dihedral (float, fixed op order).
 * n1 = (o0-e0)x(o0-e1), n2 = (o1-e1)x(o1-e0), E = e1-e0.
 * theta = atan2( ((n2 x n1).E)/|E| , n1.n2 ); flat consistently wound = 0. */
typedef struct { float E[3], n1[3], n2[3], n1sq, n2sq, elen; int ok; } dih_terms;

static void dih_compute(const float* e0, const float* e1, const float* o0, const float* o1,
                        dih_terms* d) {
    float a[3], b[3];
    v3sub(e1, e0, d->E);
    v3sub(o0, e0, a);
    v3sub(o0, e1, b);
    v3cross(a, b, d->n1);
    v3sub(o1, e1, a);
    v3sub(o1, e0, b);
    v3cross(a, b, d->n2);
    d->n1sq = v3dot(d->n1, d->n1);
    d->n2sq = v3dot(d->n2, d->n2);
    d->elen = sqrtf(v3dot(d->E, d->E));
    d->ok = (d->n1sq > 1e-24f) && (d->n2sq > 1e-24f) && (d->elen > 1e-12f);
}

static float dih_angle(const dih_terms* d) {
    float c[3];
    v3cross(d->n2, d->n1, c);
    float sn = v3dot(c, d->E) / d->elen;
    float cs = v3dot(d->n1, d->n2);
    return orc_atan2f(sn, cs);
}

float orc_dihedral(const float* e0, const float* e1, const float* o0, const float* o1) {
    dih_terms d;
    dih_compute(e0, e1, o0, o1, &d);
    if (!d.ok) return 0.0f;
    return dih_angle(&d);
}

/* Gradient (u_e0, u_e1, u_o0, u_o1) in float, widened to double for callers. */
static void dih_grad(const float* e0, const float* e1, const float* o0, const float* o1,
                     const dih_terms* d, float* u /* 12 */) {
    float t[3];
    float s1 = d->elen / d->n1sq;
    float s2 = d->elen / d->n2sq;
    /* projections of the wing vertices on the hinge, divided by |E| */
    v3sub(o0, e1, t);
    float a1 = (v3dot(t, d->E) / d->elen) / d->n1sq;
    v3sub(o1, e1, t);
    float a2 = (v3dot(t, d->E) / d->elen) / d->n2sq;
    v3sub(o0, e0, t);
    float b1 = (v3dot(t, d->E) / d->elen) / d->n1sq;
    v3sub(o1, e0, t);
    float b2 = (v3dot(t, d->E) / d->elen) / d->n2sq;
    for (int k = 0; k < 3; k++) {
        u[6 + k] = s1 * d->n1[k];                       /* o0 */
        u[9 + k] = s2 * d->n2[k];                       /* o1 */
        u[0 + k] = a1 * d->n1[k] + a2 * d->n2[k];       /* e0 */
        u[3 + k] = -(b1 * d->n1[k]) - b2 * d->n2[k];    /* e1 */
    }
}

void orc_dihedral_grad(const float* e0, const float* e1, const float* o0, const float* o1,
                       double* g) {
    dih_terms d;
    float u[12];
    dih_compute(e0, e1, o0, o1, &d);
    if (!d.ok) {
        for (int k = 0; k < 12; k++) g[k] = 0.0;
        return;
    }
    dih_grad(e0, e1, o0, o1, &d, u);
    for (int k = 0; k < 12; k++) g[k] = u[k];
}

/* FP64 twin of the angle and gradient (used by the FD gradient test and as
 * the FP64 oracle when stating FP32 error). */
void orc_dihedral_grad_f64(const double* e0, const double* e1, const double* o0,
                           const double* o1, double* theta, double* g) {
    double E[3], n1[3], n2[3], a[3], b[3];
    for (int k = 0; k < 3; k++) { E[k] = e1[k] - e0[k]; a[k] = o0[k] - e0[k]; b[k] = o0[k] - e1[k]; }
    n1[0] = a[1] * b[2] - a[2] * b[1]; n1[1] = a[2] * b[0] - a[0] * b[2]; n1[2] = a[0] * b[1] - a[1] * b[0];
    for (int k = 0; k < 3; k++) { a[k] = o1[k] - e1[k]; b[k] = o1[k] - e0[k]; }
    n2[0] = a[1] * b[2] - a[2] * b[1]; n2[1] = a[2] * b[0] - a[0] * b[2]; n2[2] = a[0] * b[1] - a[1] * b[0];
    double n1sq = n1[0] * n1[0] + n1[1] * n1[1] + n1[2] * n1[2];
    double n2sq = n2[0] * n2[0] + n2[1] * n2[1] + n2[2] * n2[2];
    double el = sqrt(E[0] * E[0] + E[1] * E[1] + E[2] * E[2]);
    double c[3] = {n2[1] * n1[2] - n2[2] * n1[1], n2[2] * n1[0] - n2[0] * n1[2], n2[0] * n1[1] - n2[1] * n1[0]};
    double sn = (c[0] * E[0] + c[1] * E[1] + c[2] * E[2]) / el;
    double cs = n1[0] * n2[0] + n1[1] * n2[1] + n1[2] * n2[2];
    *theta = atan2(sn, cs);
    double pe0[3], pe1[3], pf0[3], pf1[3];
    for (int k = 0; k < 3; k++) { pe0[k] = o0[k] - e1[k]; pe1[k] = o1[k] - e1[k]; pf0[k] = o0[k] - e0[k]; pf1[k] = o1[k] - e0[k]; }
    double a1 = (pe0[0] * E[0] + pe0[1] * E[1] + pe0[2] * E[2]) / el / n1sq;
    double a2 = (pe1[0] * E[0] + pe1[1] * E[1] + pe1[2] * E[2]) / el / n2sq;
    double b1 = (pf0[0] * E[0] + pf0[1] * E[1] + pf0[2] * E[2]) / el / n1sq;
    double b2 = (pf1[0] * E[0] + pf1[1] * E[1] + pf1[2] * E[2]) / el / n2sq;
    for (int k = 0; k < 3; k++) {
        g[6 + k] = el / n1sq * n1[k];
        g[9 + k] = el / n2sq * n2[k];
        g[0 + k] = a1 * n1[k] + a2 * n2[k];
        g[3 + k] = -b1 * n1[k] - b2 * n2[k];
    }
}

/* ------------------------------------------------------------------------ */
/* Adjacency: ConstraintPBD::GenePoint2PrimsMap / GenePrim2PrimsMap          */
/* PBD/PBD_cuda.cu:155-202.  For prim p: for each of its points in order,    */
/* every other prim on that point in ascending prim id; duplicates kept.     */
/* Returns the total neighbour count; fills the CSR when arrays are given.   */
/* ------------------------------------------------------------------------ */
int orc_prim_adjacency(const int* prim_list, int np, const int* indices, int nv,
                       int* adj_start_num, int* adj) {
    int* cnt = (int*)calloc((size_t)nv + 1, sizeof(int));
    for (int p = 0; p < np; p++)
        for (int k = 0; k < prim_list[2 * p + 1]; k++) cnt[indices[prim_list[2 * p] + k] + 1]++;
    for (int v = 0; v < nv; v++) cnt[v + 1] += cnt[v];
    int* fill = (int*)malloc(sizeof(int) * ((size_t)nv + 1));
    memcpy(fill, cnt, sizeof(int) * ((size_t)nv + 1));
    int* p2p = (int*)malloc(sizeof(int) * ((size_t)cnt[nv] + 1));
    for (int p = 0; p < np; p++)
        for (int k = 0; k < prim_list[2 * p + 1]; k++) p2p[fill[indices[prim_list[2 * p] + k]]++] = p;
    int total = 0;
    for (int p = 0; p < np; p++) {
        int start = total;
        for (int k = 0; k < prim_list[2 * p + 1]; k++) {
            int v = indices[prim_list[2 * p] + k];
            for (int j = cnt[v]; j < cnt[v + 1]; j++) {
                if (p2p[j] == p) continue;
                if (adj) adj[total] = p2p[j];
                total++;
            }
        }
        if (adj_start_num) { adj_start_num[2 * p] = start; adj_start_num[2 * p + 1] = total - start; }
    }
    free(cnt); free(fill); free(p2p);
    return total;
}

/* ------------------------------------------------------------------------ */
/* Colouring with the reference GPU semantics: AssignColorsGPU               */
/* (PBD/PBD_cuda.cu:313-372) then ResolveConflictsGPU (:374-431), each a     */
/* bulk-synchronous pass (every read sees the previous pass's arrays), run   */
/* until no prim is uncoloured.  color/prd_color must start at -1.           */
/* Returns the number of rounds used, or -1 if max_rounds was not enough.    */
/* ------------------------------------------------------------------------ */
int orc_color(const int* prim_list, int np, const int* adj_start_num, const int* adj,
              int max_rounds, int* color, int* prd_color) {
    for (int round = 0; round < max_rounds; round++) {
        int uncolored = 0;
        for (int p = 0; p < np; p++) if (color[p] < 0) uncolored++;
        if (uncolored == 0) return round;
        /* assign */
        for (int p = 0; p < np; p++) {
            if (color[p] >= 0) continue;
            int nidx = adj_start_num[2 * p], nlast = nidx + adj_start_num[2 * p + 1];
            int c = -1, offset = 0;
            while (c < 0) {
                uint64_t forbidden = 0;
                for (int i = nidx; i < nlast; ++i) {
                    int nc = color[adj[i]] - offset;
                    if (nc >= 0 && nc < 64) forbidden |= (1ull << nc);
                }
                if (forbidden ^ ~0ull) {
                    uint64_t x = (~forbidden) & (forbidden + 1);
                    c = offset;
                    while (x > 1) { x /= 2; c++; }
                } else {
                    offset += 64;
                }
            }
            prd_color[p] = c;
        }
        /* resolve (reads prd_color only; writes color of the prim itself) */
        for (int p = 0; p < np; p++) {
            if (color[p] >= 0) continue;
            int nidx = adj_start_num[2 * p], nlast = nidx + adj_start_num[2 * p + 1];
            int c = prd_color[p], conflict = 0, npt = prim_list[2 * p + 1];
            for (int i = nidx; i < nlast; ++i) {
                int n = adj[i];
                if (prd_color[n] == c) {
                    int nnpt = prim_list[2 * n + 1];
                    if (nnpt > npt || (nnpt == npt && n < p)) { conflict = 1; break; }
                }
            }
            if (!conflict) color[p] = c;
        }
    }
    for (int p = 0; p < np; p++) if (color[p] < 0) return -1;
    return max_rounds;
}

/* ConstraintPBD::SortEdgesColors (stable sort_by_key, PBD/PBD_cuda.cu:546-587)
 * + EvalWorksets (run-length (start,count), :589-625; also correct for np==1).
 * Colours are 0..C-1 (first-fit leaves no gaps).  Returns C. */
int orc_color_sort(const int* color, int np, int* sorted_prim, int* worksets) {
    int C = 0;
    for (int p = 0; p < np; p++) if (color[p] + 1 > C) C = color[p] + 1;
    int* cnt = (int*)calloc((size_t)C + 1, sizeof(int));
    for (int p = 0; p < np; p++) cnt[color[p] + 1]++;
    for (int c = 0; c < C; c++) cnt[c + 1] += cnt[c];
    for (int c = 0; c < C; c++) { worksets[2 * c] = cnt[c]; worksets[2 * c + 1] = cnt[c + 1] - cnt[c]; }
    for (int p = 0; p < np; p++) sorted_prim[cnt[color[p]]++] = p;
    free(cnt);
    return C;
}

/* ------------------------------------------------------------------------ */
/* Substep                                                                    */
/* ------------------------------------------------------------------------ */

/* AdvectGPUKernel  PBD/PBD_cuda.cu:949-979:
 *   v += g*dt*m ; v *= powf(damping, dt) ; p = x + v*dt
 * with m -> w (w = 0 pins the vertex).  Optional wind drag (new):
 *   v += ((wind - v) * (drag*dt)) * w, applied before damping. */
void orc_advect(const orc_params* prm, int nv, const float* w, const float* x, float* v, float* p) {
    float gdt[3] = {prm->gravity[0] * prm->dt, prm->gravity[1] * prm->dt, prm->gravity[2] * prm->dt};
    float kd = prm->wind_drag * prm->dt;
    for (int i = 0; i < nv; i++) {
        for (int k = 0; k < 3; k++) {
            float vk = v[3 * i + k] + gdt[k] * w[i];
            if (prm->wind_drag != 0.0f) vk = vk + ((prm->wind[k] - vk) * kd) * w[i];
            vk = vk * prm->damp_factor;
            v[3 * i + k] = vk;
            p[3 * i + k] = x[3 * i + k] + vk * prm->dt;
        }
    }
}

/* Distance projection, the reference's CPU form (projectConstraintCPU,
 * PBD/PBD_cuda.cu:1105-1124):
 *   r = normalize(p0 - p1) ; dp0 = -w0/(w0+w1) * (d-L) * r ; dp1 = +w1/(w0+w1) * (d-L) * r
 * scaled by k as the GPU kernel does (PBD/PBD_cuda.cu:1282-1287), with
 * d = sqrtf(|p0-p1|^2) and normalize = v * (1/sqrtf(v.v)) (host helper_math,
 * PBD/helper_math.h:62-65,1309-1313).  This is algebraically the GPU kernel's
 * dp0 = -w0/(w0+w1)*(d-L)*v/d (PBD/PBD_cuda.cu:1264-1281) with one division
 * instead of six.  Guards (new): skip when w0+w1 == 0 or d == 0 (the
 * reference divides by zero there). */
void orc_project_distance(const int* ids, float rest, float k, const float* w, float* p) {
    float* p0 = p + 3 * ids[0];
    float* p1 = p + 3 * ids[1];
    float w0 = w[ids[0]], w1 = w[ids[1]];
    float ws = w0 + w1;
    if (ws == 0.0f) return;
    float vx = p0[0] - p1[0], vy = p0[1] - p1[1], vz = p0[2] - p1[2];
    float d = sqrtf(vx * vx + vy * vy + vz * vz);
    if (d == 0.0f) return;
    float inv = 1.0f / d;
    float rx = vx * inv, ry = vy * inv, rz = vz * inv;
    float c = d - rest;
    float s0 = (-w0 / ws) * c;
    float s1 = (w1 / ws) * c;
    p0[0] = p0[0] + (s0 * rx) * k; p0[1] = p0[1] + (s0 * ry) * k; p0[2] = p0[2] + (s0 * rz) * k;
    p1[0] = p1[0] + (s1 * rx) * k; p1[1] = p1[1] + (s1 * ry) * k; p1[2] = p1[2] + (s1 * rz) * k;
}

/* Bending projection (new; PBD, Mueller et al. 2007 sec. 3.3 generic step
 * dp_i = -k * w_i * C / sum_j(w_j |grad_j C|^2) * grad_i C with C = theta - theta0). */
void orc_project_bending(const int* ids, float rest, float k, const float* w, float* p) {
    float* q[4] = {p + 3 * ids[0], p + 3 * ids[1], p + 3 * ids[2], p + 3 * ids[3]};
    float wq[4] = {w[ids[0]], w[ids[1]], w[ids[2]], w[ids[3]]};
    dih_terms d;
    dih_compute(q[0], q[1], q[2], q[3], &d);
    if (!d.ok) return;
    float th = dih_angle(&d);
    float u[12];
    dih_grad(q[0], q[1], q[2], q[3], &d, u);
    float den = 0.0f;
    for (int j = 0; j < 4; j++) den = den + wq[j] * v3dot(u + 3 * j, u + 3 * j);
    if (!(den > 1e-20f)) return;
    float s = ((th - rest) / den) * k;
    for (int j = 0; j < 4; j++) {
        float f = s * wq[j];
        q[j][0] = q[j][0] - f * u[3 * j + 0];
        q[j][1] = q[j][1] - f * u[3 * j + 1];
        q[j][2] = q[j][2] - f * u[3 * j + 2];
    }
}

/* One Gauss-Seidel sweep in colour order: the reference GPU loop
 * (PBD/PBD_cuda.cu:1326-1401) visits colours in workset order and, inside a
 * colour, prims that share no vertex, so a sequential pass over sortedPrimId
 * performs the same operations. */
void orc_sweep(orc_cloth* c, int sweep_index) {
    int ki = sweep_index < c->k_len ? sweep_index : c->k_len - 1;
    float kd = c->k_dist[ki], kb = c->k_bend ? c->k_bend[ki] : 1.0f;
    for (int col = 0; col < c->ncolors; col++) {
        int s = c->worksets[2 * col], n = c->worksets[2 * col + 1];
        for (int j = s; j < s + n; j++) {
            int prim = c->sorted[j];
            int st = c->prim_list[2 * prim], cnt = c->prim_list[2 * prim + 1];
            if (cnt == 2) orc_project_distance(c->indices + st, c->rest[prim], kd, c->w, c->p);
            else if (cnt == 4) orc_project_bending(c->indices + st, c->rest[prim], kb, c->w, c->p);
        }
    }
}

/* SolverPBD::ColliWithShpGrd / ColliderSphere / GenerateMoveVectorSphere
 * PBD/PBD_cuda.cu:1519-1573 with a parameterised centre and radius:
 *   if |p-c| - r <= 0.001 : p += (p-c)/|p-c| * (r - |p-c|)
 * Pinned vertices (w == 0) are left alone (new).  Box (new): a point inside
 * the axis-aligned box is pushed out through the nearest face. */
void orc_colliders(const orc_params* prm, int nv, const float* w, float* p) {
    for (int i = 0; i < nv; i++) {
        if (w[i] == 0.0f) continue;
        float* q = p + 3 * i;
        if (prm->sphere_on) {
            float d = orc_distance(q, prm->sphere_c);
            if (!(d - prm->sphere_r > 0.001f) && d > 0.0f) {
                float md = prm->sphere_r - d;
                float dx = (q[0] - prm->sphere_c[0]) / d;
                float dy = (q[1] - prm->sphere_c[1]) / d;
                float dz = (q[2] - prm->sphere_c[2]) / d;
                q[0] = q[0] + dx * md;
                q[1] = q[1] + dy * md;
                q[2] = q[2] + dz * md;
            }
        }
        if (prm->box_on) {
            int inside = 1;
            for (int k = 0; k < 3; k++) inside &= (q[k] > prm->box_lo[k]) & (q[k] < prm->box_hi[k]);
            if (inside) {
                int bk = 0, bs = 0;
                float best = q[0] - prm->box_lo[0];
                for (int k = 0; k < 3; k++) {
                    float dl = q[k] - prm->box_lo[k], dh = prm->box_hi[k] - q[k];
                    if (dl < best) { best = dl; bk = k; bs = 0; }
                    if (dh < best) { best = dh; bk = k; bs = 1; }
                }
                q[bk] = bs ? prm->box_hi[bk] : prm->box_lo[bk];
            }
        }
    }
}

/* IntegrationGPUKernel  PBD/PBD_cuda.cu:1467-1482: v = (p-x)/dt ; x = p */
void orc_integrate(int nv, float dt, float* x, float* v, const float* p) {
    for (int i = 0; i < 3 * nv; i++) {
        v[i] = (p[i] - x[i]) / dt;
        x[i] = p[i];
    }
}

/* Convergence residual (new): over distance prims, r = |d - L| / L with d
 * computed in FP64 from the FP32 positions.  Returns max r and sum r^2. */
void orc_residual(const orc_cloth* c, const float* p, double* max_rel, double* sum_sq, int* count) {
    double mx = 0.0, ss = 0.0;
    int n = 0;
    for (int prim = 0; prim < c->np; prim++) {
        if (c->prim_list[2 * prim + 1] != 2) continue;
        const int* ids = c->indices + c->prim_list[2 * prim];
        const float* a = p + 3 * ids[0];
        const float* b = p + 3 * ids[1];
        double dx = (double)a[0] - (double)b[0];
        double dy = (double)a[1] - (double)b[1];
        double dz = (double)a[2] - (double)b[2];
        double d = sqrt(dx * dx + dy * dy + dz * dz);
        double L = (double)c->rest[prim];
        double r = fabs(d - L) / L;
        if (r > mx) mx = r;
        ss += r * r;
        n++;
    }
    *max_rel = mx;
    if (sum_sq) *sum_sq = ss;
    if (count) *count = n;
}

/* Reference kernel stiffness rule k = 1 - powf(1 - s, 1/(iter+1))
 * (PBD/PBD_cuda.cu:1282), tabulated per sweep index. */
void orc_stiffness_table_ref(float s, int n, float* k) {
    for (int i = 0; i < n; i++) k[i] = 1 - powf(1 - s, 1.0 / (i + 1));
}

/* Mueller et al. 2007 rule k' = 1 - (1 - s)^(1/I), constant over sweeps. */
void orc_stiffness_table_const(float s, int iterations, int n, float* k) {
    float kk = 1 - powf(1 - s, 1.0 / iterations);
    for (int i = 0; i < n; i++) k[i] = kk;
}

/* One substep: advect (done by the caller so it can insert a warm start),
 * then sweeps in chunks of check_every until the residual after a chunk is
 * <= tol or `cap` sweeps have run, colliders after every sweep (as
 * projectConstraintCPU applies ColliWithShpGrd, PBD/PBD_cuda.cu:1125), then
 * integrate.  tol < 0 means exactly `cap` sweeps.  Returns sweeps run. */
int orc_substep(orc_cloth* c, const orc_params* prm, int cap, int check_every, double tol,
                double* final_residual) {
    int sweeps = 0;
    double r = -1.0;
    if (check_every < 1) check_every = 1;
    while (sweeps < cap) {
        orc_sweep(c, sweeps);
        if (prm->sphere_on || prm->box_on) orc_colliders(prm, c->nv, c->w, c->p);
        sweeps++;
        if (tol >= 0.0 && sweeps % check_every == 0) {
            orc_residual(c, c->p, &r, NULL, NULL);
            if (r <= tol) break;
        }
    }
    if (final_residual) {
        if (r < 0.0 || sweeps % check_every != 0) orc_residual(c, c->p, &r, NULL, NULL);
        *final_residual = r;
    }
    orc_integrate(c->nv, prm->dt, c->x, c->v, c->p);
    return sweeps;
}

/* The plain substep's sweeps and integrate on nthreads host threads (new:
 * the reference's CPU path, projectConstraintCPU PBD/PBD_cuda.cu:1090-1127,
 * is single-threaded).  The prims of one colour share no vertex, so a
 * colour's prims are split over the threads and each colour ends at the
 * threads' barrier; colliders and integrate are per vertex.  The result
 * equals orc_substep's (fixed sweeps) bit for bit.  Advect beforehand, as
 * for orc_substep.  Returns the sweeps run. */
int orc_substep_omp(orc_cloth* c, const orc_params* prm, int cap, int nthreads) {
    if (nthreads < 1) nthreads = 1;
    const int coll = prm->sphere_on || prm->box_on;
    const int nv = c->nv;
#pragma omp parallel num_threads(nthreads)
    {
        for (int sw = 0; sw < cap; sw++) {
            const int ki = sw < c->k_len ? sw : c->k_len - 1;
            const float kd = c->k_dist[ki], kb = c->k_bend ? c->k_bend[ki] : 1.0f;
            for (int col = 0; col < c->ncolors; col++) {
                const int s = c->worksets[2 * col], n = c->worksets[2 * col + 1];
#pragma omp for schedule(static)
                for (int j = s; j < s + n; j++) {
                    const int prim = c->sorted[j];
                    const int st = c->prim_list[2 * prim], cnt = c->prim_list[2 * prim + 1];
                    if (cnt == 2) orc_project_distance(c->indices + st, c->rest[prim], kd, c->w, c->p);
                    else if (cnt == 4) orc_project_bending(c->indices + st, c->rest[prim], kb, c->w, c->p);
                }
            }
            if (coll) {
#pragma omp for schedule(static)
                for (int i = 0; i < nv; i++) orc_colliders(prm, 1, c->w + i, c->p + 3 * i);
            }
        }
#pragma omp for schedule(static)
        for (int i = 0; i < nv; i++) orc_integrate(1, prm->dt, c->x + 3 * i, c->v + 3 * i, c->p + 3 * i);
    }
    return cap;
}

/* ------------------------------------------------------------------------ */
/* Jacobi sweep with constraint averaging (new).  The reference declares     */
/* SolverType::JACOBI (PBD/PBD_Basic.cuh:46-51) and never implements it; this */
/* is the averaged Jacobi of Macklin et al. 2014 ("Unified particle physics  */
/* for real-time applications", sec. 4.2): every prim is projected alone on  */
/* the sweep's input positions (the Gauss-Seidel projection arithmetic, on a */
/* copy), and a free vertex moves by omega / n times the sum of the          */
/* corrections its n prims give it, summed in ascending prim id.  Pinned     */
/* vertices (w = 0) keep their position.                                     */
/* ------------------------------------------------------------------------ */
void orc_jacobi_sweep(const orc_cloth* c, int sweep_index, float omega, const float* pin, float* pout) {
    int ki = sweep_index < c->k_len ? sweep_index : c->k_len - 1;
    float kd = c->k_dist[ki], kb = c->k_bend ? c->k_bend[ki] : 1.0f;
    int nv = c->nv;
    int* start = (int*)calloc((size_t)nv + 1, sizeof(int));
    for (int p = 0; p < c->np; p++)
        for (int j = 0; j < c->prim_list[2 * p + 1]; j++) start[c->indices[c->prim_list[2 * p] + j] + 1]++;
    for (int i = 0; i < nv; i++) start[i + 1] += start[i];
    int* fill = (int*)malloc(sizeof(int) * (size_t)(nv > 0 ? nv : 1));
    int* inc = (int*)malloc(sizeof(int) * (size_t)(start[nv] > 0 ? start[nv] : 1));
    for (int i = 0; i < nv; i++) fill[i] = start[i];
    for (int p = 0; p < c->np; p++) /* ascending prim id per vertex */
        for (int j = 0; j < c->prim_list[2 * p + 1]; j++) inc[fill[c->indices[c->prim_list[2 * p] + j]]++] = p;
    static const int ids[4] = {0, 1, 2, 3};
    for (int i = 0; i < nv; i++) {
        const float* pi = pin + 3 * i;
        float* po = pout + 3 * i;
        po[0] = pi[0]; po[1] = pi[1]; po[2] = pi[2];
        if (c->w[i] == 0.0f) continue;
        float sum[3] = {0.0f, 0.0f, 0.0f};
        for (int e = start[i]; e < start[i + 1]; e++) {
            int p = inc[e];
            int st = c->prim_list[2 * p], n = c->prim_list[2 * p + 1];
            float loc[12], wl[4];
            int role = 0;
            for (int j = 0; j < n; j++) {
                int v = c->indices[st + j];
                loc[3 * j] = pin[3 * v]; loc[3 * j + 1] = pin[3 * v + 1]; loc[3 * j + 2] = pin[3 * v + 2];
                wl[j] = c->w[v];
                if (v == i) role = j;
            }
            if (n == 2) orc_project_distance(ids, c->rest[p], kd, wl, loc);
            else orc_project_bending(ids, c->rest[p], kb, wl, loc);
            for (int k = 0; k < 3; k++) sum[k] = sum[k] + (loc[3 * role + k] - pi[k]);
        }
        int m = start[i + 1] - start[i];
        if (m > 0) {
            float f = omega / (float)m;
            for (int k = 0; k < 3; k++) po[k] = pi[k] + f * sum[k];
        }
    }
    free(inc);
    free(fill);
    free(start);
}

/* Sweeps of the Jacobi variant with the analytic colliders after each, then
 * integrate: orc_substep with orc_jacobi_sweep in place of orc_sweep. */
int orc_substep_jacobi(orc_cloth* c, const orc_params* prm, int cap, float omega, int check_every, double tol,
                       double* final_residual) {
    int sweeps = 0;
    double r = -1.0;
    float* tmp = (float*)malloc(sizeof(float) * 3 * (size_t)(c->nv > 0 ? c->nv : 1));
    if (check_every < 1) check_every = 1;
    while (sweeps < cap) {
        orc_jacobi_sweep(c, sweeps, omega, c->p, tmp);
        memcpy(c->p, tmp, sizeof(float) * 3 * (size_t)c->nv);
        if (prm->sphere_on || prm->box_on) orc_colliders(prm, c->nv, c->w, c->p);
        sweeps++;
        if (tol >= 0.0 && sweeps % check_every == 0) {
            orc_residual(c, c->p, &r, NULL, NULL);
            if (r <= tol) break;
        }
    }
    free(tmp);
    if (final_residual) {
        if (r < 0.0 || sweeps % check_every != 0) orc_residual(c, c->p, &r, NULL, NULL);
        *final_residual = r;
    }
    orc_integrate(c->nv, prm->dt, c->x, c->v, c->p);
    return sweeps;
}

/* ------------------------------------------------------------------------ */
/* FP64 twin of the plain substep (new): the same steps as orc_advect,       */
/* orc_sweep, orc_colliders and orc_integrate, on the same topology, colour   */
/* order, rest values (FP32, widened) and stiffness tables, in double, so the */
/* FP32 solver's rounding error can be stated against it.  Bending uses the  */
/* FP64 angle and gradient of orc_dihedral_grad_f64 (atan2 from libm).        */
/* ------------------------------------------------------------------------ */
static void orc64_project_distance(const int* ids, double rest, double k, const float* w, double* p) {
    double* p0 = p + 3 * ids[0];
    double* p1 = p + 3 * ids[1];
    double w0 = w[ids[0]], w1 = w[ids[1]], ws = w0 + w1;
    if (ws == 0.0) return;
    double vx = p0[0] - p1[0], vy = p0[1] - p1[1], vz = p0[2] - p1[2];
    double d = sqrt(vx * vx + vy * vy + vz * vz);
    if (d == 0.0) return;
    double c = d - rest;
    double s0 = (-w0 / ws) * c / d, s1 = (w1 / ws) * c / d;
    p0[0] += s0 * vx * k; p0[1] += s0 * vy * k; p0[2] += s0 * vz * k;
    p1[0] += s1 * vx * k; p1[1] += s1 * vy * k; p1[2] += s1 * vz * k;
}

static void orc64_project_bending(const int* ids, double rest, double k, const float* w, double* p) {
    double* q[4] = {p + 3 * ids[0], p + 3 * ids[1], p + 3 * ids[2], p + 3 * ids[3]};
    double th, u[12];
    orc_dihedral_grad_f64(q[0], q[1], q[2], q[3], &th, u);
    double den = 0.0;
    for (int j = 0; j < 4; j++)
        den += w[ids[j]] * (u[3 * j] * u[3 * j] + u[3 * j + 1] * u[3 * j + 1] + u[3 * j + 2] * u[3 * j + 2]);
    if (!(den > 1e-20)) return;
    double s = (th - rest) / den * k;
    for (int j = 0; j < 4; j++)
        for (int a = 0; a < 3; a++) q[j][a] -= s * w[ids[j]] * u[3 * j + a];
}

static void orc64_colliders(const orc_params* prm, int nv, const float* w, double* p) {
    for (int i = 0; i < nv; i++) {
        if (w[i] == 0.0f) continue;
        double* q = p + 3 * i;
        if (prm->sphere_on) {
            double dx = q[0] - prm->sphere_c[0], dy = q[1] - prm->sphere_c[1], dz = q[2] - prm->sphere_c[2];
            double d = sqrt(dx * dx + dy * dy + dz * dz);
            if (!(d - prm->sphere_r > 0.001) && d > 0.0) {
                double md = prm->sphere_r - d;
                q[0] += dx / d * md; q[1] += dy / d * md; q[2] += dz / d * md;
            }
        }
        if (prm->box_on) {
            int inside = 1;
            for (int k = 0; k < 3; k++) inside &= (q[k] > prm->box_lo[k]) & (q[k] < prm->box_hi[k]);
            if (inside) {
                int bk = 0, bs = 0;
                double best = q[0] - prm->box_lo[0];
                for (int k = 0; k < 3; k++) {
                    double dl = q[k] - prm->box_lo[k], dh = prm->box_hi[k] - q[k];
                    if (dl < best) { best = dl; bk = k; bs = 0; }
                    if (dh < best) { best = dh; bk = k; bs = 1; }
                }
                q[bk] = bs ? prm->box_hi[bk] : prm->box_lo[bk];
            }
        }
    }
}

/* `substeps` FP64 substeps of `sweeps` sweeps on x, v (3 nv doubles each);
 * p is scratch of the same size.  Topology, weights, rest values and
 * stiffness tables come from c. */
void orc64_substeps(const orc_cloth* c, const orc_params* prm, double* x, double* v, double* p, int sweeps,
                    int substeps) {
    const double dt = prm->dt, damp = prm->damp_factor, kd_w = (double)prm->wind_drag * dt;
    for (int s = 0; s < substeps; s++) {
        for (int i = 0; i < c->nv; i++)
            for (int k = 0; k < 3; k++) {
                double vk = v[3 * i + k] + (double)prm->gravity[k] * dt * c->w[i];
                if (prm->wind_drag != 0.0f) vk += ((double)prm->wind[k] - vk) * kd_w * c->w[i];
                vk *= damp;
                v[3 * i + k] = vk;
                p[3 * i + k] = x[3 * i + k] + vk * dt;
            }
        for (int sw = 0; sw < sweeps; sw++) {
            int ki = sw < c->k_len ? sw : c->k_len - 1;
            double kd = c->k_dist[ki], kb = c->k_bend ? c->k_bend[ki] : 1.0;
            for (int col = 0; col < c->ncolors; col++) {
                int st0 = c->worksets[2 * col], n = c->worksets[2 * col + 1];
                for (int j = st0; j < st0 + n; j++) {
                    int prim = c->sorted[j];
                    int st = c->prim_list[2 * prim], cnt = c->prim_list[2 * prim + 1];
                    if (cnt == 2) orc64_project_distance(c->indices + st, c->rest[prim], kd, c->w, p);
                    else if (cnt == 4) orc64_project_bending(c->indices + st, c->rest[prim], kb, c->w, p);
                }
            }
            if (prm->sphere_on || prm->box_on) orc64_colliders(prm, c->nv, c->w, p);
        }
        for (int i = 0; i < 3 * c->nv; i++) {
            v[i] = (p[i] - x[i]) / dt;
            x[i] = p[i];
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Broadphase                                                                 */
/* ------------------------------------------------------------------------ */

/* SpatialHashSystem::InitSH  PBD/SpatialHashSystem.cpp:88-98 */
void orc_grid_init(orc_grid* g, const float centre[3], const int size[3], const float cell[3]) {
    for (int k = 0; k < 3; k++) {
        float shift = (float)((unsigned)size[k]) * cell[k] / 2;
        g->start[k] = centre[k] - shift;
        g->cell[k] = cell[k];
        g->size[k] = size[k];
    }
}

/* evalTriCentroids  PBD/SpatialHashSystem.cpp:276-298 */
void orc_tri_centroids(const float* pos, const int* tris, int nt, float* cen) {
    for (int t = 0; t < nt; t++) {
        const float* a = pos + 3 * tris[3 * t];
        const float* b = pos + 3 * tris[3 * t + 1];
        const float* c = pos + 3 * tris[3 * t + 2];
        for (int k = 0; k < 3; k++) cen[3 * t + k] = (a[k] + b[k] + c[k]) / 3;
    }
}

/* calcGridPosCPU  PBD/SpatialHashSystem.cpp:391-399 */
void orc_cell_coords(const orc_grid* g, const float* c, int* gp) {
    for (int k = 0; k < 3; k++) gp[k] = (int)floorf((c[k] - g->start[k]) / g->cell[k]);
}

/* Dense mode: calcGridHashCPU + outsideGrid (PBD/SpatialHashSystem.cpp:401-413,
 * 123-130).  Hashed mode (new): Teschner et al. 2003 prime hash into a
 * power-of-two table, uint32 arithmetic; every cell has a slot. */
int orc_cell_key(const orc_grid* g, const int* gp) {
    if (g->mode == 0) {
        if (gp[0] < 0 || gp[1] < 0 || gp[2] < 0 || gp[0] >= g->size[0] || gp[1] >= g->size[1] ||
            gp[2] >= g->size[2])
            return -1;
        return (int)(((unsigned)gp[2] * (unsigned)g->size[1]) * (unsigned)g->size[0] +
                     (unsigned)gp[1] * (unsigned)g->size[0] + (unsigned)gp[0]);
    }
    unsigned h = ((unsigned)gp[0] * 73856093u) ^ ((unsigned)gp[1] * 19349663u) ^
                 ((unsigned)gp[2] * 83492791u);
    return (int)(h & ((1u << g->table_bits) - 1u));
}

int orc_num_cells(const orc_grid* g) {
    return g->mode == 0 ? g->size[0] * g->size[1] * g->size[2] : (1 << g->table_bits);
}

/* calcHashCPU  PBD/SpatialHashSystem.cpp:300-324 */
void orc_tri_keys(const orc_grid* g, const float* pos, const int* tris, int nt, int* keys, int* ids) {
    float c[3];
    int gp[3];
    for (int t = 0; t < nt; t++) {
        const float* a = pos + 3 * tris[3 * t];
        const float* b = pos + 3 * tris[3 * t + 1];
        const float* d = pos + 3 * tris[3 * t + 2];
        for (int k = 0; k < 3; k++) c[k] = (a[k] + b[k] + d[k]) / 3;
        orc_cell_coords(g, c, gp);
        keys[t] = orc_cell_key(g, gp);
        ids[t] = t;
    }
}

/* sortTrianglesCPU: thrust::sort_by_key on host vectors (stable, -1 first)
 * PBD/SpatialHashSystem.cpp:326-336.  Stable merge sort. */
static void msort(int* k, int* v, int* tk, int* tv, int n) {
    if (n < 2) return;
    int h = n / 2;
    msort(k, v, tk, tv, h);
    msort(k + h, v + h, tk, tv, n - h);
    int i = 0, j = h, o = 0;
    while (i < h && j < n) {
        if (k[j] < k[i]) { tk[o] = k[j]; tv[o++] = v[j++]; }
        else { tk[o] = k[i]; tv[o++] = v[i++]; }
    }
    while (i < h) { tk[o] = k[i]; tv[o++] = v[i++]; }
    while (j < n) { tk[o] = k[j]; tv[o++] = v[j++]; }
    memcpy(k, tk, sizeof(int) * (size_t)n);
    memcpy(v, tv, sizeof(int) * (size_t)n);
}

void orc_stable_sort_pairs(int* keys, int* vals, int n) {
    int* tk = (int*)malloc(sizeof(int) * ((size_t)n + 1));
    int* tv = (int*)malloc(sizeof(int) * ((size_t)n + 1));
    msort(keys, vals, tk, tv, n);
    free(tk); free(tv);
}

/* Cell ranges [start, end] (inclusive) of each key in the sorted key array.
 * The reference findCellStartCPU (PBD/SpatialHashSystem.cpp:349-389) misses
 * some starts and never resets; this computes the correct ranges (every
 * cell reset to -1, start/end = first/last sorted index of the key). */
void orc_cell_ranges(const int* sorted_keys, int nt, int ncells, int* cstart, int* cend) {
    for (int c = 0; c < ncells; c++) { cstart[c] = -1; cend[c] = -1; }
    for (int i = 0; i < nt; i++) {
        int k = sorted_keys[i];
        if (k < 0) continue;
        if (i == 0 || sorted_keys[i - 1] != k) cstart[k] = i;
        if (i == nt - 1 || sorted_keys[i + 1] != k) cend[k] = i;
    }
}

static int cmp_i64(const void* a, const void* b) {
    int64_t x = *(const int64_t*)a, y = *(const int64_t*)b;
    return x < y ? -1 : x > y;
}

/* Candidate vertex-face pairs: FindNeighbors (PBD/SpatialHashSystem.cpp:134-187)
 * as consumed by CCD_SH_Extended (PBD/CCD.cpp:1541-1586): for each triangle t
 * whose centroid cell is valid, each vertex v of t, each triangle f != t
 * stored in one of the 27 cells around t's cell (z, y, x order, skipping
 * cells outside a dense grid) -> candidate (v, f).  In hashed mode a face
 * counts only in the slot of its own centroid cell (cells that merely share
 * the slot are skipped), which makes the set independent of the table size;
 * in dense mode every slot is one cell and the check never fails.  Returned
 * as the sorted
 * set of keys (v << 32 | f) with pairs where v is a vertex of f removed (the
 * narrowphase rejects them first, PBD/CCD.cpp:1737-1738).  Returns the set
 * size (writes at most `cap` keys). */
long long orc_candidate_pairs(const orc_grid* g, const float* pos, const int* tris, int nt,
                              int nv, long long cap, int64_t* out) {
    (void)nv;
    int* keys = (int*)malloc(sizeof(int) * ((size_t)nt + 1));
    int* ids = (int*)malloc(sizeof(int) * ((size_t)nt + 1));
    orc_tri_keys(g, pos, tris, nt, keys, ids);
    int* sk = (int*)malloc(sizeof(int) * ((size_t)nt + 1));
    memcpy(sk, keys, sizeof(int) * (size_t)nt);
    orc_stable_sort_pairs(sk, ids, nt);
    int ncells = orc_num_cells(g);
    int* cs = (int*)malloc(sizeof(int) * (size_t)ncells);
    int* ce = (int*)malloc(sizeof(int) * (size_t)ncells);
    orc_cell_ranges(sk, nt, ncells, cs, ce);
    size_t bufcap = 1024, n = 0;
    int64_t* buf = (int64_t*)malloc(sizeof(int64_t) * bufcap);
    float c[3];
    int gp[3], q[3];
    int* tc = (int*)malloc(sizeof(int) * 3 * ((size_t)nt + 1));
    for (int t = 0; t < nt; t++) {
        const int* tv = tris + 3 * t;
        for (int k = 0; k < 3; k++)
            c[k] = (pos[3 * tv[0] + k] + pos[3 * tv[1] + k] + pos[3 * tv[2] + k]) / 3;
        orc_cell_coords(g, c, tc + 3 * t);
    }
    for (int t = 0; t < nt; t++) {
        if (keys[t] < 0) continue;
        const int* tv = tris + 3 * t;
        for (int k = 0; k < 3; k++)
            c[k] = (pos[3 * tv[0] + k] + pos[3 * tv[1] + k] + pos[3 * tv[2] + k]) / 3;
        orc_cell_coords(g, c, gp);
        for (int z = -1; z <= 1; z++)
            for (int y = -1; y <= 1; y++)
                for (int x = -1; x <= 1; x++) {
                    q[0] = gp[0] + x; q[1] = gp[1] + y; q[2] = gp[2] + z;
                    int h = orc_cell_key(g, q);
                    if (h < 0 || cs[h] < 0) continue;
                    for (int i = cs[h]; i <= ce[h]; i++) {
                        int f = ids[i];
                        if (f == t) continue;
                        if (tc[3 * f] != q[0] || tc[3 * f + 1] != q[1] || tc[3 * f + 2] != q[2]) continue;
                        const int* fv = tris + 3 * f;
                        for (int k = 0; k < 3; k++) {
                            int v = tv[k];
                            if (v == fv[0] || v == fv[1] || v == fv[2]) continue;
                            if (n == bufcap) { bufcap *= 2; buf = (int64_t*)realloc(buf, sizeof(int64_t) * bufcap); }
                            buf[n++] = ((int64_t)v << 32) | (int64_t)(uint32_t)f;
                        }
                    }
                }
    }
    qsort(buf, n, sizeof(int64_t), cmp_i64);
    size_t u = 0;
    for (size_t i = 0; i < n; i++) if (i == 0 || buf[i] != buf[i - 1]) buf[u++] = buf[i];
    if (out) for (size_t i = 0; i < u && (long long)i < cap; i++) out[i] = buf[i];
    free(buf); free(keys); free(ids); free(sk); free(cs); free(ce); free(tc);
    return (long long)u;
}

/* ------------------------------------------------------------------------ */
/* CCD: cubic solver PBD/CCD.cpp:48-113, ExtendedVFTest PBD/CCD.cpp:1733-1808 */
/* ------------------------------------------------------------------------ */
int orc_solve_quadratic(double a, double b, double c, double x[2]) {
    double d = b * b - 4 * a * c;
    if (d < 0) {
        x[0] = -b / (2 * a);
        return 0;
    }
    double q = -(b + (b < 0 ? -1.0 : 1.0) * sqrt(d)) / 2;
    int i = 0;
    if (fabs(a) > 1e-12 * fabs(q)) x[i++] = q / a;
    if (fabs(q) > 1e-12 * fabs(c)) x[i++] = c / q;
    if (i == 2 && x[0] > x[1]) { double t = x[0]; x[0] = x[1]; x[1] = t; }
    return i;
}

double orc_newtons_method(double a, double b, double c, double d, double x0, int init_dir) {
    if (init_dir != 0) {
        double y0 = d + x0 * (c + x0 * (b + x0 * a));
        double ddy0 = 2 * b + x0 * (6 * a);
        x0 += init_dir * sqrt(fabs(2 * y0 / ddy0));
    }
    for (int iter = 0; iter < 100; iter++) {
        double y = d + x0 * (c + x0 * (b + x0 * a));
        double dy = c + x0 * (2 * b + x0 * 3 * a);
        if (dy == 0) return x0;
        double x1 = x0 - y / dy;
        if (fabs(x0 - x1) < 1e-6) return x0;
        x0 = x1;
    }
    return x0;
}

/* The reference leaves xc[0] unset when solve_quadratic returns 0 roots from
 * its d >= 0 branch (PBD/CCD.cpp:51-54,99-113), which happens when a = b = 0
 * (the cubic is linear, e.g. a vertex translating against a static face).
 * Newton's method on a linear function converges from any finite start, so
 * the unset start is defined as 0. */
int orc_solve_cubic(double a, double b, double c, double d, double x[3]) {
    double xc[2] = {NAN, NAN};
    int ncrit = orc_solve_quadratic(3 * a, 2 * b, c, xc);
    if (ncrit == 0) {
        if (xc[0] != xc[0]) xc[0] = 0.0;
        x[0] = orc_newtons_method(a, b, c, d, xc[0], 0);
        return 1;
    } else if (ncrit == 1) {
        return orc_solve_quadratic(b, c, d, x);
    } else {
        double yc[2] = {d + xc[0] * (c + xc[0] * (b + xc[0] * a)), d + xc[1] * (c + xc[1] * (b + xc[1] * a))};
        int i = 0;
        if (yc[0] * a >= 0) x[i++] = orc_newtons_method(a, b, c, d, xc[0], -1);
        if (yc[0] * yc[1] <= 0) {
            int closer = fabs(yc[0]) < fabs(yc[1]) ? 0 : 1;
            x[i++] = orc_newtons_method(a, b, c, d, xc[closer], closer == 0 ? 1 : -1);
        }
        if (yc[1] * a <= 0) x[i++] = orc_newtons_method(a, b, c, d, xc[1], 1);
        return i;
    }
}

/* BarycentricCoord  PBD/CCD.cpp:115-130 */
void orc_barycentric(const float* p, const float* a, const float* b, const float* c, float* uvw) {
    float AB[3], AC[3], AP[3];
    v3sub(a, c, AB);
    v3sub(b, c, AC);
    v3sub(p, c, AP);
    float d00 = v3dot(AC, AC), d01 = v3dot(AC, AB), d02 = v3dot(AC, AP);
    float d11 = v3dot(AB, AB), d12 = v3dot(AB, AP);
    float inv = 1 / (d00 * d11 - d01 * d01);
    float v = (d11 * d02 - d01 * d12) * inv;
    float u = (d00 * d12 - d01 * d02) * inv;
    float w = 1 - u - v;
    uvw[0] = u; uvw[1] = v; uvw[2] = w;
}

static float stp(const float* u, const float* v, const float* w) {
    float c[3];
    v3cross(v, w, c);
    return v3dot(u, c);
}

/* pos(x, p, t) = x + t*(p - x) with t narrowed to float (helper_math has no
 * double overload)  PBD/CCD.cpp:8-11 */
static void lerp3(const float* x, const float* p, double t, float* r) {
    float tf = (float)t;
    for (int k = 0; k < 3; k++) r[k] = x[k] + tf * (p[k] - x[k]);
}

static void dsort(double* t, int n) {
    for (int i = 1; i < n; i++) {
        double v = t[i];
        int j = i - 1;
        while (j >= 0 && t[j] > v) { t[j + 1] = t[j]; j--; }
        t[j + 1] = v;
    }
}

/* ExtendedVFTest (free function)  PBD/CCD.cpp:1733-1808.  Vertex x0->p0
 * against face (x1,x2,x3)->(p1,p2,p3).  The caller rejects v in f.
 * Non-finite cubic roots are skipped (new). */
int orc_vf_ccd(const float* x0, const float* x1, const float* x2, const float* x3,
               const float* p0, const float* p1, const float* p2, const float* p3,
               float thickness, float* t_out, float* n_out, double* w_out) {
    float v0[3], X1[3], X2[3], X3[3], V1[3], V2[3], V3[3], tmp[3];
    v3sub(p0, x0, v0);
    v3sub(x1, x0, X1); v3sub(x2, x0, X2); v3sub(x3, x0, X3);
    v3sub(p1, x1, tmp); v3sub(tmp, v0, V1);
    v3sub(p2, x2, tmp); v3sub(tmp, v0, V2);
    v3sub(p3, x3, tmp); v3sub(tmp, v0, V3);
    double a0 = stp(X1, X2, X3);
    double a1 = (double)(stp(V1, X2, X3) + stp(X1, V2, X3) + stp(X1, X2, V3));
    double a2 = (double)(stp(X1, V2, V3) + stp(V1, X2, V3) + stp(V1, V2, X3));
    double a3 = stp(V1, V2, V3);
    double t[5];
    int nsol = orc_solve_cubic(a3, a2, a1, a0, t);
    int m = 0;
    for (int i = 0; i < nsol; i++) if (isfinite(t[i])) t[m++] = t[i];
    t[m] = 1;
    t[m + 1] = 0;
    dsort(t, m + 2);
    for (int i = 0; i < m + 2; i++) {
        if (t[i] < 0 || t[i] > 1) continue;
        float c0[3], c1[3], c2[3], c3[3], e1[3], e2[3], cr[3], cn[3], rel[3], wgt[3];
        lerp3(x0, p0, t[i], c0); lerp3(x1, p1, t[i], c1);
        lerp3(x2, p2, t[i], c2); lerp3(x3, p3, t[i], c3);
        v3sub(p2, p1, e1); v3sub(p3, p1, e2); v3cross(e1, e2, cr);
        float nrm[3];
        v3normalize(cr, nrm);
        v3sub(c2, c1, e1); v3sub(c3, c1, e2); v3cross(e1, e2, cr);
        v3normalize(cr, cn);
        v3sub(c0, c1, rel);
        double dist = v3dot(rel, cn);
        orc_barycentric(c0, c1, c2, c3, wgt);
        /* IO::Distance (powf form) in the reference; sqrtf form here, the
         * same value up to a rare last-bit difference (pinned by test) */
        float ed1 = orc_distance(c1, c2), ed2 = orc_distance(c1, c3), ed3 = orc_distance(c2, c3);
        float eps = thickness / ((ed1 + ed2 + ed3) / 3.0);
        int in1 = (wgt[0] <= (1 + eps)) && (wgt[0] >= (-eps));
        int in2 = (wgt[1] <= (1 + eps)) && (wgt[1] >= (-eps));
        int in3 = (wgt[2] <= (1 + eps)) && (wgt[2] >= (-eps));
        if ((fabs(dist) - 2.0f * thickness < -1e-6) && in1 && in2 && in3) {
            if (t_out) *t_out = (float)t[i];
            if (n_out) { n_out[0] = nrm[0]; n_out[1] = nrm[1]; n_out[2] = nrm[2]; }
            if (w_out) { w_out[0] = 1; w_out[1] = wgt[0]; w_out[2] = wgt[1]; w_out[3] = wgt[2]; }
            return 1;
        }
    }
    return 0;
}

/* NarrowVFCCDTest  PBD/CCD.cpp:1810-1872: the vertex crosses the face during
 * the substep.  Same coplanarity cubic as orc_vf_ccd; only its roots are
 * tried (the reference appends t = 1 but its loop stops before it, :1822-1825),
 * and at the first root in [0, 1] where the vertex is within 1e-6 of the
 * face's plane with barycentric weights in [-eps, 1 + eps] (eps = thickness /
 * mean edge of the end-state face, __USE_EPSILON = 1, CCD_Basic.h:10) the
 * pair crosses.  Unset or non-finite roots are skipped, as in orc_vf_ccd. */
int orc_vf_cross(const float* x0, const float* x1, const float* x2, const float* x3,
                 const float* p0, const float* p1, const float* p2, const float* p3, float thickness) {
    float v0[3], X1[3], X2[3], X3[3], V1[3], V2[3], V3[3], tmp[3];
    v3sub(p0, x0, v0);
    v3sub(x1, x0, X1); v3sub(x2, x0, X2); v3sub(x3, x0, X3);
    v3sub(p1, x1, tmp); v3sub(tmp, v0, V1);
    v3sub(p2, x2, tmp); v3sub(tmp, v0, V2);
    v3sub(p3, x3, tmp); v3sub(tmp, v0, V3);
    double a0 = stp(X1, X2, X3);
    double a1 = (double)(stp(V1, X2, X3) + stp(X1, V2, X3) + stp(X1, X2, V3));
    double a2 = (double)(stp(X1, V2, V3) + stp(V1, X2, V3) + stp(V1, V2, X3));
    double a3 = stp(V1, V2, V3);
    double t[3];
    int nsol = orc_solve_cubic(a3, a2, a1, a0, t);
    for (int i = 0; i < nsol; i++) {
        if (!isfinite(t[i]) || t[i] < 0 || t[i] > 1) continue;
        float c0[3], c1[3], c2[3], c3[3], e1[3], e2[3], cr[3], cn[3], rel[3], wgt[3];
        lerp3(x0, p0, t[i], c0); lerp3(x1, p1, t[i], c1);
        lerp3(x2, p2, t[i], c2); lerp3(x3, p3, t[i], c3);
        v3sub(c2, c1, e1); v3sub(c3, c1, e2); v3cross(e1, e2, cr);
        v3normalize(cr, cn);
        v3sub(c0, c1, rel);
        double d = v3dot(rel, cn);
        orc_barycentric(c0, c1, c2, c3, wgt);
        float ed1 = orc_distance(p1, p2), ed2 = orc_distance(p1, p3), ed3 = orc_distance(p2, p3);
        float eps = thickness / ((ed1 + ed2 + ed3) / 3.0);
        int in1 = (wgt[0] <= (1 + eps)) && (wgt[0] >= (-eps));
        int in2 = (wgt[1] <= (1 + eps)) && (wgt[1] >= (-eps));
        int in3 = (wgt[2] <= (1 + eps)) && (wgt[2] >= (-eps));
        if (fabs(d) < 1e-6 && in1 && in2 && in3) return 1;
    }
    return 0;
}

/* The reference's post-resolve check, CCD_SH_Narrow (PBD/CCD.cpp:1588-1652,
 * run after CollisionResolveNew at PositionBasedDynamics.cpp:437-441): the
 * (v, f) pairs, v not in f, that cross between x and p by orc_vf_cross.
 * Candidates are the swept-box set, which holds every crossing pair (the
 * vertex is on the face at the crossing, so the two boxes over the substep
 * meet).  With w, pairs whose four vertices are all static (w == 0) are
 * skipped: static meshes with duplicated vertices at their creases (a box
 * made of six grids) touch themselves without any motion.  Writes up to cap
 * keys (v << 32 | f), ascending; returns the count. */
long long orc_crossings(const float* x, const float* p, const int* tris, int nt, int nv, float thickness,
                        const float* w, long long cap, int64_t* out) {
    long long nk = orc_swept_pairs(x, p, tris, nt, nv, thickness, 0, 0, NULL);
    int64_t* keys = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nk + 1));
    orc_swept_pairs(x, p, tris, nt, nv, thickness, 0, nk, keys);
    long long n = 0;
    for (long long i = 0; i < nk; i++) {
        int v = (int)(keys[i] >> 32), f = (int)(keys[i] & 0xffffffff);
        const int* fv = tris + 3 * f;
        if (w && w[v] == 0.0f && w[fv[0]] == 0.0f && w[fv[1]] == 0.0f && w[fv[2]] == 0.0f) continue;
        if (orc_vf_cross(x + 3 * v, x + 3 * fv[0], x + 3 * fv[1], x + 3 * fv[2], p + 3 * v, p + 3 * fv[0],
                         p + 3 * fv[1], p + 3 * fv[2], thickness)) {
            if (out && n < cap) out[n] = keys[i];
            n++;
        }
    }
    free(keys);
    return n;
}

/* VFResolveTest  PBD/CCD.cpp:1996-2032.  min_n2 is the degenerate-face
 * bound on |(p2 - p1) x (p3 - p1)|^2, fixed at 1e-6 in the reference
 * (PBD/CCD.cpp:2008). */
int orc_vf_resolve_test(const float* x0, const float* x1, const float* x2, const float* x3,
                        const float* p0, const float* p1, const float* p2, const float* p3,
                        float thickness, float min_n2) {
    float e1[3], e2[3], cr[3], n0[3], pn[3], rel[3];
    v3sub(x2, x1, e1); v3sub(x3, x1, e2); v3cross(e1, e2, cr); v3normalize(cr, n0);
    v3sub(x0, x1, rel);
    int dir = v3dot(rel, n0) >= 0;
    v3sub(p2, p1, e1); v3sub(p3, p1, e2); v3cross(e1, e2, cr);
    if (v3dot(cr, cr) <= min_n2) return 0;
    v3normalize(cr, pn);
    float s = dir ? 1.0f : -1.0f;
    float sn[3] = {pn[0] * s, pn[1] * s, pn[2] * s};
    v3sub(p0, p1, rel);
    float depth = v3dot(rel, sn) - 2.0 * thickness;
    return depth < -1e-6;
}

/* VFResolveNew  PBD/CCD.cpp:2034-2090 (position correction only) */
void orc_vf_resolve(const float* x0, const float* x1, const float* x2, const float* x3,
                    float* p0, float* p1, float* p2, float* p3, float thickness) {
    float e1[3], e2[3], cr[3], n0[3], pn[3], rel[3], proj[3];
    v3sub(x2, x1, e1); v3sub(x3, x1, e2); v3cross(e1, e2, cr); v3normalize(cr, n0);
    v3sub(x0, x1, rel);
    int dir = v3dot(rel, n0) >= 0;
    v3sub(p2, p1, e1); v3sub(p3, p1, e2); v3cross(e1, e2, cr); v3normalize(cr, pn);
    float s = dir ? 1.0f : -1.0f;
    float sn[3] = {pn[0] * s, pn[1] * s, pn[2] * s};
    for (int k = 0; k < 3; k++) proj[k] = p1[k] + ((pn[k] * s) * 2.0f) * thickness;
    v3sub(proj, p0, rel);
    float depth = v3dot(rel, sn) * 0.5;
    float wgt[3];
    orc_barycentric(p0, p1, p2, p3, wgt);
    float sw = wgt[0] * wgt[0] + wgt[1] * wgt[1] + wgt[2] * wgt[2];
    for (int k = 0; k < 3; k++) {
        float vfix = depth * sn[k];
        float ffix = depth * (pn[k] * -s);
        p0[k] = p0[k] + vfix;
        p1[k] = p1[k] + (ffix * wgt[0]) / sw;
        p2[k] = p2[k] + (ffix * wgt[1]) / sw;
        p3[k] = p3[k] + (ffix * wgt[2]) / sw;
    }
}

/* Narrowphase over the candidate set (ExtendedVFTest per candidate, as
 * CCD_SH_Extended does, PBD/CCD.cpp:1541-1586).  The reference appends
 * contacts in discovery order and de-duplicates by linear scan (IsDuplicated,
 * PBD/CCD.cpp:132-147); here the candidate set is already unique and sorted,
 * so the contact list is in ascending (v, f) order.  x = start-of-substep
 * positions (meshTopol.posBuffer), p = current predicted positions. */
long long orc_ccd_detect(const orc_grid* g, const float* x, const float* p, const int* tris, int nt,
                         int nv, float thickness, long long cap, orc_contact* out) {
    long long nk = orc_candidate_pairs(g, p, tris, nt, nv, 0, NULL);
    int64_t* keys = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nk + 1));
    orc_candidate_pairs(g, p, tris, nt, nv, nk, keys);
    long long nc = 0;
    for (long long i = 0; i < nk; i++) {
        int v = (int)(keys[i] >> 32), f = (int)(keys[i] & 0xffffffff);
        const int* fv = tris + 3 * f;
        float t, n[3];
        double w[4];
        if (orc_vf_ccd(x + 3 * v, x + 3 * fv[0], x + 3 * fv[1], x + 3 * fv[2], p + 3 * v, p + 3 * fv[0],
                       p + 3 * fv[1], p + 3 * fv[2], thickness, &t, n, w)) {
            if (out && nc < cap) {
                orc_contact* c = out + nc;
                c->v = v; c->f = f; c->t = t;
                c->n[0] = n[0]; c->n[1] = n[1]; c->n[2] = n[2];
                for (int k = 0; k < 4; k++) c->w[k] = w[k];
            }
            nc++;
        }
    }
    free(keys);
    return nc;
}

/* Swept-box candidate set (new, no reference counterpart): (v, f) with v a
 * vertex of some triangle, v not in f, and AABB(x_v, p_v) meeting the AABB
 * of f's six start and end vertex positions widened per side by
 * 4 thickness + 1e-6 |bound|.  Brute force over all (v, f), ascending keys
 * (v << 32 | f).  mode 2 drops the box test: every (v, f), v not in f, as
 * CCD_N2 (PBD/CCD.cpp:1040-1098) visits them. */
static float widen_lo(float v, float m) { return v - (m + 1e-6f * fabsf(v)); }
static float widen_hi(float v, float m) { return v + (m + 1e-6f * fabsf(v)); }

long long orc_swept_pairs(const float* x, const float* p, const int* tris, int nt, int nv, float thickness,
                          int all_pairs, long long cap, int64_t* out) {
    float* fb = (float*)malloc(sizeof(float) * 6 * ((size_t)nt + 1));
    char* vin = (char*)calloc((size_t)nv + 1, 1);
    const float m = 4.0f * thickness;
    for (int f = 0; f < nt; f++) {
        const int* fv = tris + 3 * f;
        for (int k = 0; k < 3; k++) {
            float lo = x[3 * fv[0] + k], hi = lo;
            for (int i = 0; i < 3; i++) {
                float a = x[3 * fv[i] + k], b = p[3 * fv[i] + k];
                lo = fminf(lo, fminf(a, b));
                hi = fmaxf(hi, fmaxf(a, b));
            }
            fb[6 * f + k] = widen_lo(lo, m);
            fb[6 * f + 3 + k] = widen_hi(hi, m);
        }
        vin[fv[0]] = vin[fv[1]] = vin[fv[2]] = 1;
    }
    long long n = 0;
    for (int v = 0; v < nv; v++) {
        if (!vin[v]) continue;
        float lo[3], hi[3];
        for (int k = 0; k < 3; k++) {
            lo[k] = fminf(x[3 * v + k], p[3 * v + k]);
            hi[k] = fmaxf(x[3 * v + k], p[3 * v + k]);
        }
        for (int f = 0; f < nt; f++) {
            const int* fv = tris + 3 * f;
            if (fv[0] == v || fv[1] == v || fv[2] == v) continue;
            if (!all_pairs) {
                const float* b = fb + 6 * f;
                if (lo[0] > b[3] || b[0] > hi[0] || lo[1] > b[4] || b[1] > hi[1] || lo[2] > b[5] || b[2] > hi[2])
                    continue;
            }
            if (out && n < cap) out[n] = ((int64_t)v << 32) | (int64_t)(uint32_t)f;
            n++;
        }
    }
    free(fb);
    free(vin);
    return n;
}

/* Narrowphase (as orc_ccd_detect) over the swept-box set, or over every pair
 * when all_pairs is set. */
long long orc_ccd_detect_swept(const float* x, const float* p, const int* tris, int nt, int nv, float thickness,
                               int all_pairs, long long cap, orc_contact* out) {
    long long nk = orc_swept_pairs(x, p, tris, nt, nv, thickness, all_pairs, 0, NULL);
    int64_t* keys = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nk + 1));
    orc_swept_pairs(x, p, tris, nt, nv, thickness, all_pairs, nk, keys);
    long long nc = 0;
    for (long long i = 0; i < nk; i++) {
        int v = (int)(keys[i] >> 32), f = (int)(keys[i] & 0xffffffff);
        const int* fv = tris + 3 * f;
        float t, n[3];
        double w[4];
        if (orc_vf_ccd(x + 3 * v, x + 3 * fv[0], x + 3 * fv[1], x + 3 * fv[2], p + 3 * v, p + 3 * fv[0],
                       p + 3 * fv[1], p + 3 * fv[2], thickness, &t, n, w)) {
            if (out && nc < cap) {
                orc_contact* c = out + nc;
                c->v = v; c->f = f; c->t = t;
                c->n[0] = n[0]; c->n[1] = n[1]; c->n[2] = n[2];
                for (int k = 0; k < 4; k++) c->w[k] = w[k];
            }
            nc++;
        }
    }
    free(keys);
    return nc;
}

/* Fixed-point accumulation unit: 2^32 per metre.  Integer sums do not depend
 * on the order of accumulation, which keeps a parallel accumulation
 * deterministic and equal to this sequential one. */
static long long to_fixed(float f) { return (long long)rint((double)f * 4294967296.0); }

/* Contact response, Jacobi form of CollisionResolveNew (PBD/CCD.cpp:1874-1994):
 * each pass evaluates VFResolveTest/VFResolveNew (PBD/CCD.cpp:1996-2090) for
 * every contact against the positions at the start of the pass, sums the
 * per-vertex fixes (fixed point) and moves every vertex by the mean of its
 * fixes.  The reference runs the same per-contact rule Gauss-Seidel style in
 * discovery order.  Static parts (w == 0): a pinned vertex or face vertex is
 * not moved, and the other side takes the full separation (twice the
 * reference's half depth). */
void orc_ccd_resolve(const orc_contact* c, long long nc, const int* tris, const float* x, float* p,
                     const float* w, float thickness, int passes, float min_n2) {
    (void)w;
    int nv = 0;
    for (long long i = 0; i < nc; i++) {
        const int* fv = tris + 3 * c[i].f;
        int m = c[i].v;
        for (int k = 0; k < 3; k++) if (fv[k] > m) m = fv[k];
        if (m + 1 > nv) nv = m + 1;
    }
    long long* acc = (long long*)malloc(sizeof(long long) * 3 * (size_t)(nv + 1));
    int* cnt = (int*)malloc(sizeof(int) * (size_t)(nv + 1));
    float* q = (float*)malloc(sizeof(float) * 12);
    for (int pass = 0; pass < passes; pass++) {
        memset(acc, 0, sizeof(long long) * 3 * (size_t)nv);
        memset(cnt, 0, sizeof(int) * (size_t)nv);
        for (long long i = 0; i < nc; i++) {
            int id[4] = {c[i].v, tris[3 * c[i].f], tris[3 * c[i].f + 1], tris[3 * c[i].f + 2]};
            for (int k = 0; k < 4; k++) {
                q[3 * k] = p[3 * id[k]]; q[3 * k + 1] = p[3 * id[k] + 1]; q[3 * k + 2] = p[3 * id[k] + 2];
            }
            const float *x0 = x + 3 * id[0], *x1 = x + 3 * id[1], *x2 = x + 3 * id[2], *x3 = x + 3 * id[3];
            if (!orc_vf_resolve_test(x0, x1, x2, x3, q, q + 3, q + 6, q + 9, thickness, min_n2)) continue;
            float before[12];
            memcpy(before, q, sizeof(before));
            orc_vf_resolve(x0, x1, x2, x3, q, q + 3, q + 6, q + 9, thickness);
            float wv = w[id[0]];
            int face_static = (w[id[1]] == 0.0f) && (w[id[2]] == 0.0f) && (w[id[3]] == 0.0f);
            for (int k = 0; k < 4; k++) {
                float sc = 1.0f;
                if (w[id[k]] == 0.0f) continue;
                if (k == 0 && face_static) sc = 2.0f;
                if (k > 0 && wv == 0.0f) sc = 2.0f;
                for (int a = 0; a < 3; a++) acc[3 * id[k] + a] += to_fixed((q[3 * k + a] - before[3 * k + a]) * sc);
                cnt[id[k]]++;
            }
        }
        for (int v = 0; v < nv; v++) {
            if (!cnt[v]) continue;
            for (int a = 0; a < 3; a++)
                p[3 * v + a] = p[3 * v + a] + (float)((double)acc[3 * v + a] / 4294967296.0 / (double)cnt[v]);
        }
    }
    free(acc); free(cnt); free(q);
}

/* ------------------------------------------------------------------------ */
/* Edge-edge CCD (new: the reference declares Contact::EE, PBD/CCD_Basic.h:14,*/
/* and leaves it unimplemented, PBD/CCD.cpp:925-928).                        */
/* ------------------------------------------------------------------------ */

/* Edges in first-sighting order (triangle by triangle, edges (v0,v1),
 * (v1,v2), (v2,v0)), endpoints (lower, higher); tedge[3t+k] = edge id of
 * triangle t's edge k.  Returns the edge count.  Quadratic search: test
 * sizes only. */
int orc_edge_table(const int* tris, int nt, int* tedge, int* edges) {
    int ne = 0;
    for (int t = 0; t < nt; t++)
        for (int k = 0; k < 3; k++) {
            int a = tris[3 * t + k], b = tris[3 * t + (k + 1) % 3];
            int lo = a < b ? a : b, hi = a < b ? b : a, id = -1;
            for (int e = 0; e < ne; e++)
                if (edges[2 * e] == lo && edges[2 * e + 1] == hi) { id = e; break; }
            if (id < 0) { id = ne++; edges[2 * id] = lo; edges[2 * id + 1] = hi; }
            tedge[3 * t + k] = id;
        }
    return ne;
}

static float clamp01f(float v) { return v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v); }

/* Closest points a0 + s (a1 - a0), b0 + u (b1 - b0), s, u in [0, 1]
 * (Ericson 2005, Real-Time Collision Detection, sec. 5.1.9). */
void orc_seg_closest(const float* a0, const float* a1, const float* b0, const float* b1, float* s_out,
                     float* u_out) {
    float d1[3], d2[3], r[3];
    v3sub(a1, a0, d1); v3sub(b1, b0, d2); v3sub(a0, b0, r);
    float a = v3dot(d1, d1), e = v3dot(d2, d2), f = v3dot(d2, r);
    float s, u;
    if (a <= 1e-30f && e <= 1e-30f) { s = 0.0f; u = 0.0f; }
    else if (a <= 1e-30f) { s = 0.0f; u = clamp01f(f / e); }
    else {
        float c = v3dot(d1, r);
        if (e <= 1e-30f) { u = 0.0f; s = clamp01f(-c / a); }
        else {
            float b = v3dot(d1, d2);
            float den = a * e - b * b;
            s = den > 0.0f ? clamp01f((b * f - c * e) / den) : 0.0f;
            u = (b * s + f) / e;
            if (u < 0.0f) { u = 0.0f; s = clamp01f(-c / a); }
            else if (u > 1.0f) { u = 1.0f; s = clamp01f((b - c) / a); }
        }
    }
    *s_out = s;
    *u_out = u;
}

static void seg_point(const float* a0, const float* a1, float s, float* r) {
    for (int k = 0; k < 3; k++) r[k] = a0[k] + s * (a1[k] - a0[k]);
}

/* Whether two edges can be in contact: not near parallel (sin^2 > 1e-12)
 * and, at the start positions X, at least one thickness apart along the
 * normal of the edges (q1-q0) x (q3-q2).  side = offset of the closest
 * points at the start, nn = the normal. */
static int ee_separable(const float* X0, const float* X1, const float* X2, const float* X3, const float* q0,
                        const float* q1, const float* q2, const float* q3, float thickness, float* side,
                        float* nn) {
    float s0, u0, pa[3], pb[3], ea[3], eb[3], cr[3];
    orc_seg_closest(X0, X1, X2, X3, &s0, &u0);
    seg_point(X0, X1, s0, pa);
    seg_point(X2, X3, u0, pb);
    v3sub(pa, pb, side);
    v3sub(q1, q0, ea); v3sub(q3, q2, eb); v3cross(ea, eb, cr);
    if (v3dot(cr, cr) <= 1e-12f * (v3dot(ea, ea) * v3dot(eb, eb))) return 0;
    v3normalize(cr, nn);
    if (!(fabsf(v3dot(side, nn)) >= thickness)) return 0;
    return 1;
}

/* Edge (x0,x1)->(p0,p1) against edge (x2,x3)->(p2,p3), for separable edges
 * (ee_separable at the end state): the coplanarity times of the four points
 * (the cubic of orc_vf_ccd with the first edge's start point as origin;
 * Provot 1997) plus 0 and 1, ascending; the first time in [0, 1] at which
 * the segments are within 2 x thickness is the contact: t, the end-state
 * edge normal (p1-p0) x (p3-p2) normalised, weights (1-s, s, 1-u, u). */
int orc_ee_ccd(const float* x0, const float* x1, const float* x2, const float* x3,
               const float* p0, const float* p1, const float* p2, const float* p3,
               float thickness, float* t_out, float* n_out, double* w_out) {
    float side[3], nrm[3];
    if (!ee_separable(x0, x1, x2, x3, p0, p1, p2, p3, thickness, side, nrm)) return 0;
    float v0[3], X1[3], X2[3], X3[3], V1[3], V2[3], V3[3], tmp[3];
    v3sub(p0, x0, v0);
    v3sub(x1, x0, X1); v3sub(x2, x0, X2); v3sub(x3, x0, X3);
    v3sub(p1, x1, tmp); v3sub(tmp, v0, V1);
    v3sub(p2, x2, tmp); v3sub(tmp, v0, V2);
    v3sub(p3, x3, tmp); v3sub(tmp, v0, V3);
    double a0 = stp(X1, X2, X3);
    double a1 = (double)(stp(V1, X2, X3) + stp(X1, V2, X3) + stp(X1, X2, V3));
    double a2 = (double)(stp(X1, V2, V3) + stp(V1, X2, V3) + stp(V1, V2, X3));
    double a3 = stp(V1, V2, V3);
    double t[5];
    int nsol = orc_solve_cubic(a3, a2, a1, a0, t);
    int m = 0;
    for (int i = 0; i < nsol; i++) if (isfinite(t[i])) t[m++] = t[i];
    t[m] = 1;
    t[m + 1] = 0;
    dsort(t, m + 2);
    for (int i = 0; i < m + 2; i++) {
        if (t[i] < 0 || t[i] > 1) continue;
        float c0[3], c1[3], c2[3], c3[3], pa[3], pb[3], s, u;
        lerp3(x0, p0, t[i], c0); lerp3(x1, p1, t[i], c1);
        lerp3(x2, p2, t[i], c2); lerp3(x3, p3, t[i], c3);
        orc_seg_closest(c0, c1, c2, c3, &s, &u);
        seg_point(c0, c1, s, pa);
        seg_point(c2, c3, u, pb);
        float d = orc_distance(pa, pb);
        if (d - 2.0f * thickness < -1e-6f) {
            if (t_out) *t_out = (float)t[i];
            if (n_out) { n_out[0] = nrm[0]; n_out[1] = nrm[1]; n_out[2] = nrm[2]; }
            if (w_out) { w_out[0] = 1.0f - s; w_out[1] = s; w_out[2] = 1.0f - u; w_out[3] = u; }
            return 1;
        }
    }
    return 0;
}

/* Edge-edge response on the current positions q (start positions X orient
 * it), for separable edges (ee_separable): the closest points are pushed
 * apart along the edge normal, oriented as their offset at the start, until
 * 2 x thickness apart, half each, spread over each edge's ends by
 * w_k / sum w^2.  d: 4 displacements (12 floats).  Returns 0 if none. */
int orc_ee_fix(const float* X0, const float* X1, const float* X2, const float* X3, const float* q0,
               const float* q1, const float* q2, const float* q3, float thickness, float* d) {
    float pa[3], pb[3], side[3], nn[3], rel[3], s, u;
    if (!ee_separable(X0, X1, X2, X3, q0, q1, q2, q3, thickness, side, nn)) return 0;
    if (v3dot(side, nn) < 0.0f) { nn[0] = -nn[0]; nn[1] = -nn[1]; nn[2] = -nn[2]; }
    orc_seg_closest(q0, q1, q2, q3, &s, &u);
    seg_point(q0, q1, s, pa);
    seg_point(q2, q3, u, pb);
    v3sub(pa, pb, rel);
    float gap = v3dot(rel, nn);
    if (!(gap - 2.0f * thickness < -1e-6f)) return 0;
    float depth = (2.0f * thickness - gap) * 0.5f;
    float w[4] = {1.0f - s, s, 1.0f - u, u};
    float sa = w[0] * w[0] + w[1] * w[1], sb = w[2] * w[2] + w[3] * w[3];
    for (int k = 0; k < 4; k++) {
        float f = k < 2 ? (depth * w[k]) / sa : -((depth * w[k]) / sb);
        d[3 * k] = nn[0] * f; d[3 * k + 1] = nn[1] * f; d[3 * k + 2] = nn[2] * f;
    }
    return 1;
}

static int cmp_pair(const void* a, const void* b) {
    const int* x = (const int*)a;
    const int* y = (const int*)b;
    if (x[0] != y[0]) return x[0] < y[0] ? -1 : 1;
    if (x[1] != y[1]) return x[1] < y[1] ? -1 : 1;
    return 0;
}

/* Whether a vertex of edge a is a mesh neighbour (joined by an edge) of a
 * vertex of edge b; off/nbr: vertex -> neighbours. */
static int ring_neighbours(const int* a, const int* b, const int* off, const int* nbr) {
    for (int i = 0; i < 2; i++)
        for (int j = off[a[i]]; j < off[a[i] + 1]; j++)
            if (nbr[j] == b[0] || nbr[j] == b[1]) return 1;
    return 0;
}

/* Edge-edge candidates by brute force over triangle pairs t < f whose
 * centroid cells are both inside the grid and at most one cell apart on
 * each axis: every edge pair sharing no vertex and with no vertex of one a
 * mesh neighbour of a vertex of the other, as (lower edge id, higher),
 * distinct and ascending.  Returns the count; writes min(count, cap). */
long long orc_ee_candidates(const orc_grid* g, const float* pos, const int* tris, int nt, const int* tedge,
                            const int* edges, long long cap, int* out) {
    int ne = 0, nv = 0;
    for (int i = 0; i < 3 * nt; i++) if (tedge[i] + 1 > ne) ne = tedge[i] + 1;
    for (int i = 0; i < 2 * ne; i++) if (edges[i] + 1 > nv) nv = edges[i] + 1;
    int* off = (int*)calloc((size_t)nv + 1, sizeof(int));
    int* nbr = (int*)malloc(sizeof(int) * (size_t)(2 * ne + 1));
    int* fill = (int*)malloc(sizeof(int) * (size_t)(nv + 1));
    for (int e = 0; e < ne; e++) { off[edges[2 * e] + 1]++; off[edges[2 * e + 1] + 1]++; }
    for (int v = 0; v < nv; v++) off[v + 1] += off[v];
    for (int v = 0; v < nv; v++) fill[v] = off[v];
    for (int e = 0; e < ne; e++) {
        nbr[fill[edges[2 * e]]++] = edges[2 * e + 1];
        nbr[fill[edges[2 * e + 1]]++] = edges[2 * e];
    }
    int* keys = (int*)malloc(sizeof(int) * ((size_t)nt + 1));
    int* ids = (int*)malloc(sizeof(int) * ((size_t)nt + 1));
    orc_tri_keys(g, pos, tris, nt, keys, ids);
    int* tc = (int*)malloc(sizeof(int) * 3 * ((size_t)nt + 1));
    float c[3];
    for (int t = 0; t < nt; t++) {
        const int* tv = tris + 3 * t;
        for (int k = 0; k < 3; k++)
            c[k] = (pos[3 * tv[0] + k] + pos[3 * tv[1] + k] + pos[3 * tv[2] + k]) / 3;
        orc_cell_coords(g, c, tc + 3 * t);
    }
    size_t bufcap = 1024, n = 0;
    int* buf = (int*)malloc(sizeof(int) * 2 * bufcap);
    for (int t = 0; t < nt; t++) {
        if (keys[t] < 0) continue;
        for (int f = t + 1; f < nt; f++) {
            if (keys[f] < 0) continue;
            int near = 1;
            for (int k = 0; k < 3; k++) near &= abs(tc[3 * t + k] - tc[3 * f + k]) <= 1;
            if (!near) continue;
            for (int kb = 0; kb < 3; kb++)
                for (int ka = 0; ka < 3; ka++) {
                    int ea = tedge[3 * t + ka], eb = tedge[3 * f + kb];
                    const int* av = edges + 2 * ea;
                    const int* bv = edges + 2 * eb;
                    if (av[0] == bv[0] || av[0] == bv[1] || av[1] == bv[0] || av[1] == bv[1]) continue;
                    if (ring_neighbours(av, bv, off, nbr)) continue;
                    if (n == bufcap) { bufcap *= 2; buf = (int*)realloc(buf, sizeof(int) * 2 * bufcap); }
                    buf[2 * n] = ea < eb ? ea : eb;
                    buf[2 * n + 1] = ea < eb ? eb : ea;
                    n++;
                }
        }
    }
    qsort(buf, n, 2 * sizeof(int), cmp_pair);
    size_t u = 0;
    for (size_t i = 0; i < n; i++)
        if (i == 0 || buf[2 * i] != buf[2 * u - 2] || buf[2 * i + 1] != buf[2 * u - 1]) {
            buf[2 * u] = buf[2 * i];
            buf[2 * u + 1] = buf[2 * i + 1];
            u++;
        }
    if (out) for (size_t i = 0; i < u && (long long)i < cap; i++) { out[2 * i] = buf[2 * i]; out[2 * i + 1] = buf[2 * i + 1]; }
    free(buf); free(keys); free(ids); free(tc); free(off); free(nbr); free(fill);
    return (long long)u;
}

/* Edge-edge detection: candidates (above) through orc_ee_ccd, in candidate
 * order; quads = (a0, a1, b0, b1). */
long long orc_ccd_detect_ee(const orc_grid* g, const float* x, const float* p, const int* tris, int nt,
                            float thickness, long long cap, int* quads, float* t, float* n, double* w) {
    int* tedge = (int*)malloc(sizeof(int) * 3 * ((size_t)nt + 1));
    int* edges = (int*)malloc(sizeof(int) * 6 * ((size_t)nt + 1));
    orc_edge_table(tris, nt, tedge, edges);
    long long np = orc_ee_candidates(g, p, tris, nt, tedge, edges, 0, NULL);
    int* cand = (int*)malloc(sizeof(int) * 2 * (size_t)(np + 1));
    orc_ee_candidates(g, p, tris, nt, tedge, edges, np, cand);
    long long nc = 0;
    for (long long i = 0; i < np; i++) {
        const int* a = edges + 2 * cand[2 * i];
        const int* b = edges + 2 * cand[2 * i + 1];
        float tt, nn[3];
        double ww[4];
        if (!orc_ee_ccd(x + 3 * a[0], x + 3 * a[1], x + 3 * b[0], x + 3 * b[1], p + 3 * a[0], p + 3 * a[1],
                        p + 3 * b[0], p + 3 * b[1], thickness, &tt, nn, ww))
            continue;
        if (nc < cap) {
            if (quads) { quads[4 * nc] = a[0]; quads[4 * nc + 1] = a[1]; quads[4 * nc + 2] = b[0]; quads[4 * nc + 3] = b[1]; }
            if (t) t[nc] = tt;
            if (n) for (int k = 0; k < 3; k++) n[3 * nc + k] = nn[k];
            if (w) for (int k = 0; k < 4; k++) w[4 * nc + k] = ww[k];
        }
        nc++;
    }
    free(cand); free(edges); free(tedge);
    return nc;
}

/* orc_ccd_resolve with edge-edge contacts (quads) summed into the same
 * fixed-point accumulators in every pass; an edge whose two ends are static
 * leaves the whole gap to the other. */
void orc_ccd_resolve_ee(const orc_contact* c, long long nc, const int* quads, long long nee, const int* tris,
                        const float* x, float* p, const float* w, int nv, float thickness, int passes,
                        float min_n2) {
    long long* acc = (long long*)malloc(sizeof(long long) * 3 * (size_t)(nv + 1));
    int* cnt = (int*)malloc(sizeof(int) * (size_t)(nv + 1));
    float q[12], d[12];
    for (int pass = 0; pass < passes; pass++) {
        memset(acc, 0, sizeof(long long) * 3 * (size_t)nv);
        memset(cnt, 0, sizeof(int) * (size_t)nv);
        for (long long i = 0; i < nc; i++) {
            int id[4] = {c[i].v, tris[3 * c[i].f], tris[3 * c[i].f + 1], tris[3 * c[i].f + 2]};
            for (int k = 0; k < 4; k++) {
                q[3 * k] = p[3 * id[k]]; q[3 * k + 1] = p[3 * id[k] + 1]; q[3 * k + 2] = p[3 * id[k] + 2];
            }
            const float *x0 = x + 3 * id[0], *x1 = x + 3 * id[1], *x2 = x + 3 * id[2], *x3 = x + 3 * id[3];
            if (!orc_vf_resolve_test(x0, x1, x2, x3, q, q + 3, q + 6, q + 9, thickness, min_n2)) continue;
            float before[12];
            memcpy(before, q, sizeof(before));
            orc_vf_resolve(x0, x1, x2, x3, q, q + 3, q + 6, q + 9, thickness);
            float wv = w[id[0]];
            int face_static = (w[id[1]] == 0.0f) && (w[id[2]] == 0.0f) && (w[id[3]] == 0.0f);
            for (int k = 0; k < 4; k++) {
                float sc = 1.0f;
                if (w[id[k]] == 0.0f) continue;
                if (k == 0 && face_static) sc = 2.0f;
                if (k > 0 && wv == 0.0f) sc = 2.0f;
                for (int a = 0; a < 3; a++) acc[3 * id[k] + a] += to_fixed((q[3 * k + a] - before[3 * k + a]) * sc);
                cnt[id[k]]++;
            }
        }
        for (long long i = 0; i < nee; i++) {
            const int* id = quads + 4 * i;
            for (int k = 0; k < 4; k++) {
                q[3 * k] = p[3 * id[k]]; q[3 * k + 1] = p[3 * id[k] + 1]; q[3 * k + 2] = p[3 * id[k] + 2];
            }
            if (!orc_ee_fix(x + 3 * id[0], x + 3 * id[1], x + 3 * id[2], x + 3 * id[3], q, q + 3, q + 6, q + 9,
                            thickness, d))
                continue;
            int a_static = w[id[0]] == 0.0f && w[id[1]] == 0.0f, b_static = w[id[2]] == 0.0f && w[id[3]] == 0.0f;
            for (int k = 0; k < 4; k++) {
                if (w[id[k]] == 0.0f) continue;
                float sc = (k < 2 ? b_static : a_static) ? 2.0f : 1.0f;
                for (int a = 0; a < 3; a++) acc[3 * id[k] + a] += to_fixed(d[3 * k + a] * sc);
                cnt[id[k]]++;
            }
        }
        for (int v = 0; v < nv; v++) {
            if (!cnt[v]) continue;
            for (int a = 0; a < 3; a++)
                p[3 * v + a] = p[3 * v + a] + (float)((double)acc[3 * v + a] / 4294967296.0 / (double)cnt[v]);
        }
    }
    free(acc); free(cnt);
}

/* ------------------------------------------------------------------------ */
/* Warm-start MLP in FP64 (new).  22 -> 64 -> 64 -> 3, GELU (tanh form,      */
/* Hendrycks & Gimpel 2016), weights row-major W[out][in].                   */
/* Loss = mean over samples of 0.5*|y - t|^2.                                */
/* ------------------------------------------------------------------------ */
void orc_gelu(double x, double* y, double* dy) {
    const double c = 0.7978845608028654; /* sqrt(2/pi) */
    double u = c * (x + 0.044715 * x * x * x);
    double th = tanh(u);
    *y = 0.5 * x * (1.0 + th);
    if (dy) *dy = 0.5 * (1.0 + th) + 0.5 * x * (1.0 - th * th) * c * (1.0 + 3.0 * 0.044715 * x * x);
}

void orc_mlp_forward(const double* W1, const double* b1, const double* W2, const double* b2,
                     const double* W3, const double* b3, const double* x, int n, double* y,
                     double* h1, double* a1, double* h2, double* a2) {
    for (int s = 0; s < n; s++) {
        const double* xi = x + (size_t)s * ORC_MLP_IN;
        double* H1 = h1 + (size_t)s * ORC_MLP_H; double* A1 = a1 + (size_t)s * ORC_MLP_H;
        double* H2 = h2 + (size_t)s * ORC_MLP_H; double* A2 = a2 + (size_t)s * ORC_MLP_H;
        for (int o = 0; o < ORC_MLP_H; o++) {
            double acc = b1[o];
            for (int i = 0; i < ORC_MLP_IN; i++) acc += W1[o * ORC_MLP_IN + i] * xi[i];
            H1[o] = acc;
            orc_gelu(acc, &A1[o], NULL);
        }
        for (int o = 0; o < ORC_MLP_H; o++) {
            double acc = b2[o];
            for (int i = 0; i < ORC_MLP_H; i++) acc += W2[o * ORC_MLP_H + i] * A1[i];
            H2[o] = acc;
            orc_gelu(acc, &A2[o], NULL);
        }
        for (int o = 0; o < ORC_MLP_OUT; o++) {
            double acc = b3[o];
            for (int i = 0; i < ORC_MLP_H; i++) acc += W3[o * ORC_MLP_H + i] * A2[i];
            y[(size_t)s * ORC_MLP_OUT + o] = acc;
        }
    }
}

double orc_mlp_loss_grad(const double* W1, const double* b1, const double* W2,
                         const double* b2, const double* W3, const double* b3,
                         const double* x, const double* t, int n, double* gW1, double* gb1,
                         double* gW2, double* gb2, double* gW3, double* gb3) {
    double *h1 = (double*)malloc(sizeof(double) * (size_t)n * ORC_MLP_H);
    double *a1 = (double*)malloc(sizeof(double) * (size_t)n * ORC_MLP_H);
    double *h2 = (double*)malloc(sizeof(double) * (size_t)n * ORC_MLP_H);
    double *a2 = (double*)malloc(sizeof(double) * (size_t)n * ORC_MLP_H);
    double *y = (double*)malloc(sizeof(double) * (size_t)n * ORC_MLP_OUT);
    orc_mlp_forward(W1, b1, W2, b2, W3, b3, x, n, y, h1, a1, h2, a2);
    memset(gW1, 0, sizeof(double) * ORC_MLP_H * ORC_MLP_IN); memset(gb1, 0, sizeof(double) * ORC_MLP_H);
    memset(gW2, 0, sizeof(double) * ORC_MLP_H * ORC_MLP_H); memset(gb2, 0, sizeof(double) * ORC_MLP_H);
    memset(gW3, 0, sizeof(double) * ORC_MLP_OUT * ORC_MLP_H); memset(gb3, 0, sizeof(double) * ORC_MLP_OUT);
    double loss = 0.0;
    double inv = 1.0 / n;
    for (int s = 0; s < n; s++) {
        double dy[ORC_MLP_OUT], d2[ORC_MLP_H], d1[ORC_MLP_H];
        const double* xi = x + (size_t)s * ORC_MLP_IN;
        double* A1 = a1 + (size_t)s * ORC_MLP_H; double* A2 = a2 + (size_t)s * ORC_MLP_H;
        double* H1 = h1 + (size_t)s * ORC_MLP_H; double* H2 = h2 + (size_t)s * ORC_MLP_H;
        for (int o = 0; o < ORC_MLP_OUT; o++) {
            double e = y[(size_t)s * ORC_MLP_OUT + o] - t[(size_t)s * ORC_MLP_OUT + o];
            loss += 0.5 * e * e * inv;
            dy[o] = e * inv;
            gb3[o] += dy[o];
            for (int i = 0; i < ORC_MLP_H; i++) gW3[o * ORC_MLP_H + i] += dy[o] * A2[i];
        }
        for (int i = 0; i < ORC_MLP_H; i++) {
            double acc = 0.0, g, gd;
            for (int o = 0; o < ORC_MLP_OUT; o++) acc += W3[o * ORC_MLP_H + i] * dy[o];
            orc_gelu(H2[i], &g, &gd);
            d2[i] = acc * gd;
            gb2[i] += d2[i];
        }
        for (int o = 0; o < ORC_MLP_H; o++)
            for (int i = 0; i < ORC_MLP_H; i++) gW2[o * ORC_MLP_H + i] += d2[o] * A1[i];
        for (int i = 0; i < ORC_MLP_H; i++) {
            double acc = 0.0, g, gd;
            for (int o = 0; o < ORC_MLP_H; o++) acc += W2[o * ORC_MLP_H + i] * d2[o];
            orc_gelu(H1[i], &g, &gd);
            d1[i] = acc * gd;
            gb1[i] += d1[i];
        }
        for (int o = 0; o < ORC_MLP_H; o++)
            for (int i = 0; i < ORC_MLP_IN; i++) gW1[o * ORC_MLP_IN + i] += d1[o] * xi[i];
    }
    free(h1); free(a1); free(h2); free(a2); free(y);
    return loss;
}

/* Adam (Kingma & Ba 2015, Algorithm 1) with bias correction. */
void orc_adam_step(double* w, const double* g, double* m, double* v, int n, double lr,
                   double b1, double b2, double eps, int step) {
    double c1 = 1.0 - pow(b1, step), c2 = 1.0 - pow(b2, step);
    for (int i = 0; i < n; i++) {
        m[i] = b1 * m[i] + (1.0 - b1) * g[i];
        v[i] = b2 * v[i] + (1.0 - b2) * g[i] * g[i];
        double mh = m[i] / c1, vh = v[i] / c2;
        w[i] -= lr * mh / (sqrt(vh) + eps);
    }
}

/* BF16 round-to-nearest-even from FP32 (NaN kept quiet). */
uint16_t orc_f32_to_bf16(float x) {
    uint32_t u;
    memcpy(&u, &x, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    uint32_t r = u + 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(r >> 16);
}

float orc_bf16_to_f32(uint16_t h) {
    uint32_t u = (uint32_t)h << 16;
    float x;
    memcpy(&x, &u, 4);
    return x;
}
