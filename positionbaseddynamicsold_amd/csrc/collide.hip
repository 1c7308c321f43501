// Spatial-hash broadphase and vertex-face CCD for gfx950.
//
// Broadphase (SpatialHashSystem, SpatialHashSystem.cpp:88-187, 276-413):
//   triangle centroid -> cell coordinates -> key (dense grid index or -1
//   outside, or a hashed table slot) -> stable radix sort of (key, id) ->
//   cell ranges [start, end] -> candidate (vertex, face) pairs from the 27
//   cells around each triangle's cell, de-duplicated by a 64-bit radix sort.
// Narrowphase: ExtendedVFTest (CCD.cpp:1733-1808) per candidate, FP64 only in
//   the coplanarity cubic / time-of-impact solve (CCD.cpp:48-113).
// Response: VFResolveTest/VFResolveNew (CCD.cpp:1996-2090) evaluated for all
//   contacts against the same positions, fixes summed per vertex in 2^-32 m
//   fixed point (integer sums are order independent, so the result does not
//   depend on atomic ordering) and applied as the mean fix per vertex.
// Paths are relative to PositionBasedDynamicsCUDA/PositionBasedDynamicsCUDA/.
#include <mutex>

#include "pbd_internal.hpp"
#include "pbd_math.hpp"

namespace pbd {

// ---------------------------------------------------------------------------
// device scan (exclusive, int64 totals kept in int)
// ---------------------------------------------------------------------------
constexpr int SB = 1024;  // elements per scan block (256 threads x 4)

__global__ __launch_bounds__(256) void k_scan_blocks(const int* __restrict__ in, int* __restrict__ out, int n,
                                                     int* __restrict__ sums) {
    __shared__ int s[256];
    int base = blockIdx.x * SB, t = threadIdx.x;
    int v[4], tot = 0;
    for (int k = 0; k < 4; k++) {
        int i = base + t * 4 + k;
        v[k] = i < n ? in[i] : 0;
        tot += v[k];
    }
    s[t] = tot;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        int a = t >= off ? s[t - off] : 0;
        __syncthreads();
        s[t] += a;
        __syncthreads();
    }
    int run = s[t] - tot;
    for (int k = 0; k < 4; k++) {
        int i = base + t * 4 + k;
        if (i < n) out[i] = run;
        run += v[k];
    }
    if (t == 255) sums[blockIdx.x] = s[255];
}

__global__ __launch_bounds__(256) void k_scan_add(int* __restrict__ out, int n, const int* __restrict__ sums) {
    int i = blockIdx.x * SB + threadIdx.x;
    int add = sums[blockIdx.x];
    for (int k = 0; k < 4; k++, i += 256)
        if (i < n) out[i] += add;
}

__global__ __launch_bounds__(1024) void k_scan_single(int* __restrict__ a, int n) {
    __shared__ int part[1024];
    int t = threadIdx.x;
    int per = (n + 1023) / 1024;
    int b = t * per, e = min(n, b + per);
    int s = 0;
    for (int i = b; i < e; i++) s += a[i];
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        int v = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int run = part[t] - s;
    for (int i = b; i < e; i++) {
        int v = a[i];
        a[i] = run;
        run += v;
    }
}

void device_scan(hipStream_t s, const int* in, int* out, int n, int* tmp) {
    if (n <= 0) return;
    int nb = (n + SB - 1) / SB;
    hipLaunchKernelGGL(k_scan_blocks, dim3(nb), dim3(256), 0, s, in, out, n, tmp);
    hipLaunchKernelGGL(k_scan_single, dim3(1), dim3(1024), 0, s, tmp, nb);
    hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(256), 0, s, out, n, tmp);
    PBD_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// broadphase
// ---------------------------------------------------------------------------
__device__ __forceinline__ int cell_key(const HashCfg& g, int x, int y, int z) {
    if (g.mode == 0) {
        if (x < 0 || y < 0 || z < 0 || x >= g.size[0] || y >= g.size[1] || z >= g.size[2]) return -1;
        return (int)(((unsigned)z * (unsigned)g.size[1]) * (unsigned)g.size[0] + (unsigned)y * (unsigned)g.size[0] +
                     (unsigned)x);
    }
    unsigned h = ((unsigned)x * 73856093u) ^ ((unsigned)y * 19349663u) ^ ((unsigned)z * 83492791u);
    return (int)(h & ((1u << g.table_bits) - 1u));
}

// A vertex position from the library's float4 records or from 3 floats per
// vertex (the C ABI's layout).
__device__ __forceinline__ float4 vload(const float4* __restrict__ p, int i) { return p[i]; }
__device__ __forceinline__ float4 vload(const float* __restrict__ p, int i) {
    return make_float4(p[3 * (size_t)i], p[3 * (size_t)i + 1], p[3 * (size_t)i + 2], 0.0f);
}

// centroid (a+b+c)/3 per component (evalTriCentroids) and its cell
// (calcGridPosCPU: floor((c - start)/cell)).
template <typename P>
__device__ __forceinline__ void tri_cell(const HashCfg& g, const P* __restrict__ p, const int* __restrict__ tris,
                                         int t, int* gp) {
    float4 a = vload(p, tris[3 * t]), b = vload(p, tris[3 * t + 1]), c = vload(p, tris[3 * t + 2]);
    float cx = (a.x + b.x + c.x) / 3, cy = (a.y + b.y + c.y) / 3, cz = (a.z + b.z + c.z) / 3;
    gp[0] = (int)floorf((cx - g.start[0]) / g.cell[0]);
    gp[1] = (int)floorf((cy - g.start[1]) / g.cell[1]);
    gp[2] = (int)floorf((cz - g.start[2]) / g.cell[2]);
}

__global__ __launch_bounds__(256) void k_tri_keys(HashCfg g, const float4* __restrict__ p,
                                                  const int* __restrict__ tris, int nt,
                                                  uint32_t* __restrict__ keys1, uint32_t* __restrict__ ids,
                                                  int* __restrict__ raw, int4* __restrict__ tcell) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    int gp[3];
    tri_cell(g, p, tris, t, gp);
    int k = cell_key(g, gp[0], gp[1], gp[2]);
    keys1[t] = (uint32_t)(k + 1);  // -1 sorts first, as the reference's signed sort
    ids[t] = (uint32_t)t;
    raw[t] = k;
    tcell[t] = make_int4(gp[0], gp[1], gp[2], 0);
}

__global__ __launch_bounds__(256) void k_fill_i32(int* __restrict__ a, int n, int v) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// Correct ranges of each key in the sorted array (the reference's
// findCellStartCPU misses some starts, SpatialHashSystem.cpp:357-387).
__global__ __launch_bounds__(256) void k_cell_ranges(const uint32_t* __restrict__ keys1, int n,
                                                     int* __restrict__ cstart, int* __restrict__ cend) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int k = (int)keys1[i] - 1;
    if (k < 0) return;
    if (i == 0 || keys1[i - 1] != keys1[i]) cstart[k] = i;
    if (i == n - 1 || keys1[i + 1] != keys1[i]) cend[k] = i;
}

// Candidate (v, f): v a vertex of triangle t whose cell is valid, f != t in
// the 27 cells around t's cell (FindNeighbors, SpatialHashSystem.cpp:134-187),
// v not a vertex of f.  A face is taken only from the slot of its own cell,
// so hashed tables give the same set at every size (other cells that share
// the slot are skipped); dense keys are one slot per cell already.
// mode 0 counts, mode 1 writes at off[t].
template <int EMIT>
__global__ __launch_bounds__(256) void k_pairs(HashCfg g, const float4* __restrict__ p,
                                               const int* __restrict__ tris, int nt,
                                               const int* __restrict__ raw, const uint32_t* __restrict__ sids,
                                               const int* __restrict__ cstart, const int* __restrict__ cend,
                                               const int4* __restrict__ tcell, int* __restrict__ cnt_or_off,
                                               uint32_t* __restrict__ ov, uint32_t* __restrict__ of) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    int n = 0, off = EMIT ? cnt_or_off[t] : 0;
    if (raw[t] >= 0) {
        int gp[3];
        tri_cell(g, p, tris, t, gp);
        int tv[3] = {tris[3 * t], tris[3 * t + 1], tris[3 * t + 2]};
        for (int z = -1; z <= 1; z++)
            for (int y = -1; y <= 1; y++)
                for (int x = -1; x <= 1; x++) {
                    const int qx = gp[0] + x, qy = gp[1] + y, qz = gp[2] + z;
                    int h = cell_key(g, qx, qy, qz);
                    if (h < 0) continue;
                    int s = cstart[h];
                    if (s < 0) continue;
                    int e = cend[h];
                    for (int i = s; i <= e; i++) {
                        int f = (int)sids[i];
                        if (f == t) continue;
                        int4 fc = tcell[f];
                        if (fc.x != qx || fc.y != qy || fc.z != qz) continue;
                        int f0 = tris[3 * f], f1 = tris[3 * f + 1], f2 = tris[3 * f + 2];
                        for (int k = 0; k < 3; k++) {
                            int v = tv[k];
                            if (v == f0 || v == f1 || v == f2) continue;
                            if (EMIT) {
                                ov[off + n] = (uint32_t)v;
                                of[off + n] = (uint32_t)f;
                            }
                            n++;
                        }
                    }
                }
    }
    if (!EMIT) cnt_or_off[t] = n;
}

constexpr int EE_RING = 40;  // LDS slots per thread for a triangle's vertices and their neighbours

// True when a vertex of edge a is a mesh neighbour of a vertex of edge b
// (vertex -> neighbour CSR voff / vnbr).
__device__ __forceinline__ bool ring_neighbours(int2 a, int2 b, const int* __restrict__ voff,
                                                const int* __restrict__ vnbr) {
    const int ends[2] = {a.x, a.y};
    for (int k = 0; k < 2; k++)
        for (int j = voff[ends[k]]; j < voff[ends[k] + 1]; j++) {
            int u = vnbr[j];
            if (u == b.x || u == b.y) return true;
        }
    return false;
}

// Edge-edge candidates (new: the reference declares Contact::EE, CCD_Basic.h:14,
// and leaves it unimplemented, CCD.cpp:925-928).  From the triangle pairs
// k_pairs visits (centroid cells within one cell, the neighbour's own cell
// checked), taken once (t < f): every edge of t against every edge of f that
// shares no vertex with it and has no vertex that is a mesh neighbour of one
// of its vertices (pairs that close in the mesh are material neighbours, not
// contacts), as (lower edge id, higher edge id).
template <int EMIT>
__global__ __launch_bounds__(256) void k_ee_pairs(HashCfg g, const float4* __restrict__ p,
                                                  const int* __restrict__ tris, int nt, const int* __restrict__ raw,
                                                  const uint32_t* __restrict__ sids, const int* __restrict__ cstart,
                                                  const int* __restrict__ cend, const int4* __restrict__ tcell,
                                                  const int* __restrict__ tedge, const int2* __restrict__ edges,
                                                  const int* __restrict__ voff, const int* __restrict__ vnbr,
                                                  int* __restrict__ cnt_or_off, uint32_t* __restrict__ oa,
                                                  uint32_t* __restrict__ ob) {
    // t's vertices and their mesh neighbours, per thread in LDS: a triangle f
    // with no vertex in this set shares no vertex with, and is outside the
    // 2-ring of, every edge of t, so its edge pairs skip the exact tests
    __shared__ int ring[256][EE_RING + 1];
    int* rt = ring[threadIdx.x];
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    int n = 0, off = EMIT ? cnt_or_off[t] : 0;
    if (raw[t] >= 0) {
        int gp[3];
        tri_cell(g, p, tris, t, gp);
        int te[3] = {tedge[3 * t], tedge[3 * t + 1], tedge[3 * t + 2]};
        int2 tv[3] = {edges[te[0]], edges[te[1]], edges[te[2]]};
        int nr = 0;
        bool exact_all = false;
        for (int k = 0; k < 3 && !exact_all; k++) {
            const int v = tris[3 * t + k];
            const int b0 = voff[v], b1 = voff[v + 1];
            if (nr + 1 + (b1 - b0) > EE_RING) {
                exact_all = true;
                break;
            }
            rt[nr++] = v;
            for (int j = b0; j < b1; j++) rt[nr++] = vnbr[j];
        }
        for (int z = -1; z <= 1; z++)
            for (int y = -1; y <= 1; y++)
                for (int x = -1; x <= 1; x++) {
                    const int qx = gp[0] + x, qy = gp[1] + y, qz = gp[2] + z;
                    int h = cell_key(g, qx, qy, qz);
                    if (h < 0) continue;
                    int s = cstart[h];
                    if (s < 0) continue;
                    int e = cend[h];
                    for (int i = s; i <= e; i++) {
                        int f = (int)sids[i];
                        if (f <= t) continue;
                        int4 fc = tcell[f];
                        if (fc.x != qx || fc.y != qy || fc.z != qz) continue;
                        bool exact = exact_all;
                        if (!exact) {
                            const int f0 = tris[3 * f], f1 = tris[3 * f + 1], f2 = tris[3 * f + 2];
                            for (int k = 0; k < nr; k++) {
                                const int u = rt[k];
                                if (u == f0 || u == f1 || u == f2) {
                                    exact = true;
                                    break;
                                }
                            }
                        }
                        for (int kb = 0; kb < 3; kb++) {
                            int eb = tedge[3 * f + kb];
                            int2 bv = edges[eb];
                            for (int ka = 0; ka < 3; ka++) {
                                int2 av = tv[ka];
                                if (exact) {
                                    if (av.x == bv.x || av.x == bv.y || av.y == bv.x || av.y == bv.y) continue;
                                    if (ring_neighbours(av, bv, voff, vnbr)) continue;
                                }
                                if (EMIT) {
                                    int ea = te[ka];
                                    oa[off + n] = (uint32_t)(ea < eb ? ea : eb);
                                    ob[off + n] = (uint32_t)(ea < eb ? eb : ea);
                                }
                                n++;
                            }
                        }
                    }
                }
    }
    if (!EMIT) cnt_or_off[t] = n;
}

__global__ __launch_bounds__(256) void k_unique_flags(const uint32_t* __restrict__ hi, const uint32_t* __restrict__ lo,
                                                      int n, int* __restrict__ flag) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    flag[i] = (i == 0 || hi[i] != hi[i - 1] || lo[i] != lo[i - 1]) ? 1 : 0;
}

__global__ __launch_bounds__(256) void k_compact_pairs(const uint32_t* __restrict__ hi, const uint32_t* __restrict__ lo,
                                                       int n, const int* __restrict__ flag,
                                                       const int* __restrict__ pos, uint32_t* __restrict__ ohi,
                                                       uint32_t* __restrict__ olo) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !flag[i]) return;
    ohi[pos[i]] = hi[i];
    olo[pos[i]] = lo[i];
}

// ---------------------------------------------------------------------------
// Swept-box broadphase (a conservative alternative to the centroid-cell
// candidate set above, which misses faces larger than a cell).  Candidate
// (v, f): v is a vertex of some triangle, v not in f, and v's box over the
// substep, AABB(x_v, p_v), meets f's box over the substep, AABB of the six
// points x_f*, p_f*, widened per side by 4 thickness plus 1e-6 |bound|.  The
// narrowphase accepts a vertex within 2 thickness of the face plane and
// inside the face's barycentric band of width thickness / mean edge, i.e.
// within sqrt(13) thickness of the face, at an interpolated position inside
// both boxes, so every pair it accepts is a candidate: the contact set
// equals that of the all-pairs test (CCD_N2, CCD.cpp:1040-1098).  The cell
// grid only accelerates the box tests and does not change the set.
// ---------------------------------------------------------------------------
constexpr int SWEPT_MAX_CELLS = 4096;  // cells one box may span

__device__ __forceinline__ float widen_lo(float v, float m) { return v - (m + 1e-6f * fabsf(v)); }
__device__ __forceinline__ float widen_hi(float v, float m) { return v + (m + 1e-6f * fabsf(v)); }

__device__ __forceinline__ int cell_coord(const HashCfg& g, float v, int k) {
    float c = floorf((v - g.start[k]) / g.cell[k]);
    c = fminf(fmaxf(c, -536870912.0f), 536870912.0f);  // keeps hi - lo + 1 within int
    int i = (int)c;
    if (g.mode == 0) i = min(max(i, 0), g.size[k] - 1);  // clamping keeps overlapping ranges overlapping
    return i;
}

struct SweptBox {
    float lo[3], hi[3];
};

__device__ __forceinline__ SweptBox face_box(const float4* __restrict__ x, const float4* __restrict__ p,
                                             const int* __restrict__ tris, int f, float margin) {
    SweptBox b;
    const float4 a[6] = {x[tris[3 * f]], x[tris[3 * f + 1]], x[tris[3 * f + 2]],
                         p[tris[3 * f]], p[tris[3 * f + 1]], p[tris[3 * f + 2]]};
    float lo[3] = {a[0].x, a[0].y, a[0].z}, hi[3] = {a[0].x, a[0].y, a[0].z};
    for (int i = 1; i < 6; i++) {
        const float c[3] = {a[i].x, a[i].y, a[i].z};
        for (int k = 0; k < 3; k++) {
            lo[k] = fminf(lo[k], c[k]);
            hi[k] = fmaxf(hi[k], c[k]);
        }
    }
    for (int k = 0; k < 3; k++) {
        b.lo[k] = widen_lo(lo[k], margin);
        b.hi[k] = widen_hi(hi[k], margin);
    }
    return b;
}

__device__ __forceinline__ SweptBox vertex_box(const float4 X, const float4 P) {
    SweptBox b;
    b.lo[0] = fminf(X.x, P.x); b.hi[0] = fmaxf(X.x, P.x);
    b.lo[1] = fminf(X.y, P.y); b.hi[1] = fmaxf(X.y, P.y);
    b.lo[2] = fminf(X.z, P.z); b.hi[2] = fmaxf(X.z, P.z);
    return b;
}

// Face boxes, their cell ranges and cell counts; err[0] = 1 if a face spans
// more than SWEPT_MAX_CELLS cells (its count is then 0).
__global__ __launch_bounds__(256) void k_swept_faces(HashCfg g, const float4* __restrict__ x, const float4* __restrict__ p,
                                                     const int* __restrict__ tris, int nt, float margin,
                                                     float4* __restrict__ fbox, int4* __restrict__ fcell,
                                                     int* __restrict__ cnt, int* __restrict__ err) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nt) return;
    const SweptBox b = face_box(x, p, tris, f, margin);
    int lo[3], hi[3];
    long long n = 1;
    for (int k = 0; k < 3; k++) {
        lo[k] = cell_coord(g, b.lo[k], k);
        hi[k] = cell_coord(g, b.hi[k], k);
        // per-axis bound first: the product of three spans of up to 2^31 overflows
        n *= (long long)min(hi[k] - lo[k] + 1, SWEPT_MAX_CELLS + 1);
    }
    fbox[2 * f] = make_float4(b.lo[0], b.lo[1], b.lo[2], 0.0f);
    fbox[2 * f + 1] = make_float4(b.hi[0], b.hi[1], b.hi[2], 0.0f);
    fcell[2 * f] = make_int4(lo[0], lo[1], lo[2], 0);
    fcell[2 * f + 1] = make_int4(hi[0], hi[1], hi[2], 0);
    if (n > SWEPT_MAX_CELLS) {
        err[0] = 1;
        n = 0;
    }
    cnt[f] = (int)n;
}

// One (cell key + 1, face) entry per cell of each face box, x fastest.
__global__ __launch_bounds__(256) void k_swept_entries(HashCfg g, int nt, const int4* __restrict__ fcell,
                                                       const int* __restrict__ off, uint32_t* __restrict__ keys,
                                                       uint32_t* __restrict__ ids) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nt) return;
    int o = off[f];
    if (off[f + 1] == o) return;
    const int4 lo = fcell[2 * f], hi = fcell[2 * f + 1];
    for (int z = lo.z; z <= hi.z; z++)
        for (int y = lo.y; y <= hi.y; y++)
            for (int xx = lo.x; xx <= hi.x; xx++) {
                keys[o] = (uint32_t)(cell_key(g, xx, y, z) + 1);
                ids[o] = (uint32_t)f;
                o++;
            }
}

__global__ __launch_bounds__(256) void k_mark_vertices(const int* __restrict__ tris, int nt, int* __restrict__ vin) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 3 * nt) vin[tris[i]] = 1;
}

// Per vertex: faces found in the cells of its box whose boxes meet it.  A
// pair is taken at the lowest cell of the two cell ranges' intersection
// only (a face stored twice in one slot is removed by the later unique
// pass).  err[1] = 1 if a vertex box spans more than SWEPT_MAX_CELLS cells.
template <int EMIT>
__global__ __launch_bounds__(256) void k_swept_pairs(HashCfg g, const float4* __restrict__ x,
                                                     const float4* __restrict__ p, const int* __restrict__ vin, int nv,
                                                     const int* __restrict__ tris, const float4* __restrict__ fbox,
                                                     const int4* __restrict__ fcell, const uint32_t* __restrict__ sids,
                                                     const int* __restrict__ cstart, const int* __restrict__ cend,
                                                     int* __restrict__ cnt_or_off, uint32_t* __restrict__ ov,
                                                     uint32_t* __restrict__ of, int* __restrict__ err) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    int n = 0;
    const int base = EMIT ? cnt_or_off[v] : 0;
    if (vin[v]) {
        const SweptBox b = vertex_box(x[v], p[v]);
        int lo[3], hi[3];
        long long span = 1;
        for (int k = 0; k < 3; k++) {
            lo[k] = cell_coord(g, b.lo[k], k);
            hi[k] = cell_coord(g, b.hi[k], k);
            span *= (long long)min(hi[k] - lo[k] + 1, SWEPT_MAX_CELLS + 1);
        }
        if (span > SWEPT_MAX_CELLS) {
            if (!EMIT) err[1] = 1;
        } else {
            for (int qz = lo[2]; qz <= hi[2]; qz++)
                for (int qy = lo[1]; qy <= hi[1]; qy++)
                    for (int qx = lo[0]; qx <= hi[0]; qx++) {
                        const int h = cell_key(g, qx, qy, qz);
                        if (h < 0) continue;
                        const int s = cstart[h];
                        if (s < 0) continue;
                        const int e = cend[h];
                        for (int i = s; i <= e; i++) {
                            const int f = (int)sids[i];
                            const int4 fl = fcell[2 * f], fh = fcell[2 * f + 1];
                            // the pair's first shared cell, and f really covers it
                            if (qx != max(lo[0], fl.x) || qy != max(lo[1], fl.y) || qz != max(lo[2], fl.z)) continue;
                            if (qx > fh.x || qy > fh.y || qz > fh.z) continue;
                            const float4 flo = fbox[2 * f], fhi = fbox[2 * f + 1];
                            if (b.lo[0] > fhi.x || flo.x > b.hi[0] || b.lo[1] > fhi.y || flo.y > b.hi[1] ||
                                b.lo[2] > fhi.z || flo.z > b.hi[2])
                                continue;
                            if (tris[3 * f] == v || tris[3 * f + 1] == v || tris[3 * f + 2] == v) continue;
                            if (EMIT) {
                                ov[base + n] = (uint32_t)v;
                                of[base + n] = (uint32_t)f;
                            }
                            n++;
                        }
                    }
        }
    }
    if (!EMIT) cnt_or_off[v] = n;
}

// ---------------------------------------------------------------------------
// narrowphase: FP64 cubic (solve_cubic / newtons_method / solve_quadratic,
// CCD.cpp:48-113), ExtendedVFTest (CCD.cpp:1733-1808)
// ---------------------------------------------------------------------------
__device__ int d_solve_quadratic(double a, double b, double c, double* x) {
    double d = b * b - 4 * a * c;
    if (d < 0) {
        x[0] = -b / (2 * a);
        return 0;
    }
    double q = -(b + (b < 0 ? -1.0 : 1.0) * sqrt(d)) / 2;
    int i = 0;
    if (fabs(a) > 1e-12 * fabs(q)) x[i++] = q / a;
    if (fabs(q) > 1e-12 * fabs(c)) x[i++] = c / q;
    if (i == 2 && x[0] > x[1]) {
        double t = x[0];
        x[0] = x[1];
        x[1] = t;
    }
    return i;
}

__device__ double d_newton(double a, double b, double c, double d, double x0, int init_dir) {
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

// When the critical-point quadratic leaves its root unset (a = b = 0: the
// cubic is linear; the reference reads an unset value there), Newton starts
// from 0, which converges to the linear root.
__device__ int d_solve_cubic(double a, double b, double c, double d, double* x) {
    double xc[2];
    xc[0] = __longlong_as_double(0x7ff8000000000000ll);
    xc[1] = xc[0];
    int ncrit = d_solve_quadratic(3 * a, 2 * b, c, xc);
    if (ncrit == 0) {
        if (xc[0] != xc[0]) xc[0] = 0.0;
        x[0] = d_newton(a, b, c, d, xc[0], 0);
        return 1;
    } else if (ncrit == 1) {
        return d_solve_quadratic(b, c, d, x);
    }
    double yc0 = d + xc[0] * (c + xc[0] * (b + xc[0] * a)), yc1 = d + xc[1] * (c + xc[1] * (b + xc[1] * a));
    int i = 0;
    if (yc0 * a >= 0) x[i++] = d_newton(a, b, c, d, xc[0], -1);
    if (yc0 * yc1 <= 0) {
        int closer = fabs(yc0) < fabs(yc1) ? 0 : 1;
        x[i++] = d_newton(a, b, c, d, xc[closer], closer == 0 ? 1 : -1);
    }
    if (yc1 * a <= 0) x[i++] = d_newton(a, b, c, d, xc[1], 1);
    return i;
}

__device__ __forceinline__ V3 v3(float4 a) { return mk(a.x, a.y, a.z); }
__device__ __forceinline__ V3 vnormalize(V3 a) {
    float inv = 1.0f / sqrtf(dot(a, a));
    return V3{a.x * inv, a.y * inv, a.z * inv};
}
__device__ __forceinline__ float stp(V3 u, V3 v, V3 w) { return dot(u, cross(v, w)); }
__device__ __forceinline__ V3 lerp(V3 x, V3 p, double t) {
    float tf = (float)t;
    return V3{x.x + tf * (p.x - x.x), x.y + tf * (p.y - x.y), x.z + tf * (p.z - x.z)};
}
// BarycentricCoord (CCD.cpp:115-130)
__device__ __forceinline__ V3 bary(V3 pp, V3 a, V3 b, V3 c) {
    V3 AB = sub(a, c), AC = sub(b, c), AP = sub(pp, c);
    float d00 = dot(AC, AC), d01 = dot(AC, AB), d02 = dot(AC, AP), d11 = dot(AB, AB), d12 = dot(AB, AP);
    float inv = 1 / (d00 * d11 - d01 * d01);
    float v = (d11 * d02 - d01 * d12) * inv;
    float u = (d00 * d12 - d01 * d02) * inv;
    float w = 1 - u - v;
    return V3{u, v, w};
}
__device__ __forceinline__ float dist3(V3 a, V3 b) {
    V3 d = sub(a, b);
    return sqrtf(d.x * d.x + d.y * d.y + d.z * d.z);
}

__device__ bool vf_ccd(V3 x0, V3 x1, V3 x2, V3 x3, V3 p0, V3 p1, V3 p2, V3 p3, float th, float* t_out, V3* n_out,
                       V3* w_out) {
    V3 v0 = sub(p0, x0);
    V3 X1 = sub(x1, x0), X2 = sub(x2, x0), X3 = sub(x3, x0);
    V3 V1 = sub(sub(p1, x1), v0), V2 = sub(sub(p2, x2), v0), V3_ = sub(sub(p3, x3), v0);
    double a0 = stp(X1, X2, X3);
    double a1 = (double)(stp(V1, X2, X3) + stp(X1, V2, X3) + stp(X1, X2, V3_));
    double a2 = (double)(stp(X1, V2, V3_) + stp(V1, X2, V3_) + stp(V1, V2, X3));
    double a3 = stp(V1, V2, V3_);
    double t[5];
    int nsol = d_solve_cubic(a3, a2, a1, a0, t);
    int m = 0;
    for (int i = 0; i < nsol; i++)
        if (isfinite(t[i])) t[m++] = t[i];
    t[m] = 1;
    t[m + 1] = 0;
    int nt = m + 2;
    for (int i = 1; i < nt; i++) {
        double v = t[i];
        int j = i - 1;
        while (j >= 0 && t[j] > v) {
            t[j + 1] = t[j];
            j--;
        }
        t[j + 1] = v;
    }
    for (int i = 0; i < nt; i++) {
        if (t[i] < 0 || t[i] > 1) continue;
        V3 c0 = lerp(x0, p0, t[i]), c1 = lerp(x1, p1, t[i]), c2 = lerp(x2, p2, t[i]), c3 = lerp(x3, p3, t[i]);
        V3 nrm = vnormalize(cross(sub(p2, p1), sub(p3, p1)));
        V3 cn = vnormalize(cross(sub(c2, c1), sub(c3, c1)));
        double dist = dot(sub(c0, c1), cn);
        V3 wg = bary(c0, c1, c2, c3);
        float ed1 = dist3(c1, c2), ed2 = dist3(c1, c3), ed3 = dist3(c2, c3);
        float eps = th / ((ed1 + ed2 + ed3) / 3.0);
        bool in1 = (wg.x <= (1 + eps)) && (wg.x >= (-eps));
        bool in2 = (wg.y <= (1 + eps)) && (wg.y >= (-eps));
        bool in3 = (wg.z <= (1 + eps)) && (wg.z >= (-eps));
        if ((fabs(dist) - 2.0f * th < -1e-6) && in1 && in2 && in3) {
            *t_out = (float)t[i];
            *n_out = nrm;
            *w_out = wg;
            return true;
        }
    }
    return false;
}

__global__ __launch_bounds__(256) void k_narrow(const uint32_t* __restrict__ cv, const uint32_t* __restrict__ cf,
                                                int n, const int* __restrict__ tris, const float4* __restrict__ x,
                                                const float4* __restrict__ p, float th, int* __restrict__ flag,
                                                float* __restrict__ ct, float* __restrict__ cn,
                                                double* __restrict__ cw) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int v = (int)cv[i], f = (int)cf[i];
    int a = tris[3 * f], b = tris[3 * f + 1], c = tris[3 * f + 2];
    float t;
    V3 nr, wg;
    bool hit = vf_ccd(v3(x[v]), v3(x[a]), v3(x[b]), v3(x[c]), v3(p[v]), v3(p[a]), v3(p[b]), v3(p[c]), th, &t, &nr, &wg);
    flag[i] = hit ? 1 : 0;
    if (hit) {
        ct[i] = t;
        cn[3 * i] = nr.x;
        cn[3 * i + 1] = nr.y;
        cn[3 * i + 2] = nr.z;
        cw[4 * i] = 1.0;
        cw[4 * i + 1] = wg.x;
        cw[4 * i + 2] = wg.y;
        cw[4 * i + 3] = wg.z;
    }
}

__global__ __launch_bounds__(256) void k_compact_contacts(int n, const int* __restrict__ flag,
                                                          const int* __restrict__ pos,
                                                          const uint32_t* __restrict__ cv,
                                                          const uint32_t* __restrict__ cf, const float* __restrict__ ct,
                                                          const float* __restrict__ cn, const double* __restrict__ cw,
                                                          int* __restrict__ ov, int* __restrict__ of,
                                                          float* __restrict__ ot, float* __restrict__ on,
                                                          double* __restrict__ ow) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !flag[i]) return;
    int o = pos[i];
    ov[o] = (int)cv[i];
    of[o] = (int)cf[i];
    ot[o] = ct[i];
    for (int k = 0; k < 3; k++) on[3 * o + k] = cn[3 * i + k];
    for (int k = 0; k < 4; k++) ow[4 * o + k] = cw[4 * i + k];
}

// ---------------------------------------------------------------------------
// edge-edge narrowphase and response (new; see k_ee_pairs)
// ---------------------------------------------------------------------------
__device__ __forceinline__ long long to_fixed(float f) { return (long long)rint((double)f * 4294967296.0); }
__device__ __forceinline__ float clamp01(float v) { return v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v); }

// Closest points a0 + s (a1 - a0), b0 + u (b1 - b0), s, u in [0, 1], of two
// segments (Ericson 2005, Real-Time Collision Detection, sec. 5.1.9).
__device__ void seg_closest(V3 a0, V3 a1, V3 b0, V3 b1, float* s_out, float* u_out) {
    V3 d1 = sub(a1, a0), d2 = sub(b1, b0), r = sub(a0, b0);
    float a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
    float s, u;
    if (a <= 1e-30f && e <= 1e-30f) {
        s = 0.0f;
        u = 0.0f;
    } else if (a <= 1e-30f) {
        s = 0.0f;
        u = clamp01(f / e);
    } else {
        float c = dot(d1, r);
        if (e <= 1e-30f) {
            u = 0.0f;
            s = clamp01(-c / a);
        } else {
            float b = dot(d1, d2);
            float den = a * e - b * b;
            s = den > 0.0f ? clamp01((b * f - c * e) / den) : 0.0f;
            u = (b * s + f) / e;
            if (u < 0.0f) {
                u = 0.0f;
                s = clamp01(-c / a);
            } else if (u > 1.0f) {
                u = 1.0f;
                s = clamp01((b - c) / a);
            }
        }
    }
    *s_out = s;
    *u_out = u;
}

__device__ __forceinline__ V3 seg_point(V3 a0, V3 a1, float s) {
    return V3{a0.x + s * (a1.x - a0.x), a0.y + s * (a1.y - a0.y), a0.z + s * (a1.z - a0.z)};
}

// Whether two edges can be in contact: not near parallel (sin^2 of the
// angle > 1e-12) and, at the start of the step, at least one thickness apart
// along the normal n of the edges (q1 - q0) x (q3 - q2).  Edges that start
// side by side (neighbours in one sheet whose spacing is below the contact
// gap) have no separation along n to restore and are not contacts.
__device__ bool ee_separable(V3 X0, V3 X1, V3 X2, V3 X3, V3 q0, V3 q1, V3 q2, V3 q3, float th, V3* side_out,
                             V3* n_out) {
    float s0, u0;
    seg_closest(X0, X1, X2, X3, &s0, &u0);
    V3 side = sub(seg_point(X0, X1, s0), seg_point(X2, X3, u0));
    V3 ea = sub(q1, q0), eb = sub(q3, q2);
    V3 cr = cross(ea, eb);
    if (dot(cr, cr) <= 1e-12f * (dot(ea, ea) * dot(eb, eb))) return false;
    V3 nn = vnormalize(cr);
    if (!(fabsf(dot(side, nn)) >= th)) return false;
    *side_out = side;
    *n_out = nn;
    return true;
}

// Edge (x0, x1) -> (p0, p1) against edge (x2, x3) -> (p2, p3) over the step,
// for separable edges (ee_separable at the end state): the times the four
// points are coplanar (the cubic of vf_ccd, with the first edge's start
// point as origin; Provot 1997) plus t = 0 and 1, ascending; the first time
// in [0, 1] where the segments come within 2 x thickness is the contact.
// Records t, the end-state edge normal (p1 - p0) x (p3 - p2) normalised and
// the weights (1 - s, s, 1 - u, u) of the closest points.
__device__ bool ee_ccd(V3 x0, V3 x1, V3 x2, V3 x3, V3 p0, V3 p1, V3 p2, V3 p3, float th, float* t_out, V3* n_out,
                       float* w_out) {
    V3 side, nrm;
    if (!ee_separable(x0, x1, x2, x3, p0, p1, p2, p3, th, &side, &nrm)) return false;
    V3 v0 = sub(p0, x0);
    V3 X1 = sub(x1, x0), X2 = sub(x2, x0), X3 = sub(x3, x0);
    V3 V1 = sub(sub(p1, x1), v0), V2 = sub(sub(p2, x2), v0), V3_ = sub(sub(p3, x3), v0);
    double a0 = stp(X1, X2, X3);
    double a1 = (double)(stp(V1, X2, X3) + stp(X1, V2, X3) + stp(X1, X2, V3_));
    double a2 = (double)(stp(X1, V2, V3_) + stp(V1, X2, V3_) + stp(V1, V2, X3));
    double a3 = stp(V1, V2, V3_);
    double t[5];
    int nsol = d_solve_cubic(a3, a2, a1, a0, t);
    int m = 0;
    for (int i = 0; i < nsol; i++)
        if (isfinite(t[i])) t[m++] = t[i];
    t[m] = 1;
    t[m + 1] = 0;
    int nt = m + 2;
    for (int i = 1; i < nt; i++) {
        double v = t[i];
        int j = i - 1;
        while (j >= 0 && t[j] > v) {
            t[j + 1] = t[j];
            j--;
        }
        t[j + 1] = v;
    }
    for (int i = 0; i < nt; i++) {
        if (t[i] < 0 || t[i] > 1) continue;
        V3 c0 = lerp(x0, p0, t[i]), c1 = lerp(x1, p1, t[i]), c2 = lerp(x2, p2, t[i]), c3 = lerp(x3, p3, t[i]);
        float s, u;
        seg_closest(c0, c1, c2, c3, &s, &u);
        float d = dist3(seg_point(c0, c1, s), seg_point(c2, c3, u));
        if (d - 2.0f * th < -1e-6f) {
            *t_out = (float)t[i];
            *n_out = nrm;
            w_out[0] = 1.0f - s;
            w_out[1] = s;
            w_out[2] = 1.0f - u;
            w_out[3] = u;
            return true;
        }
    }
    return false;
}

__global__ __launch_bounds__(256) void k_narrow_ee(const uint32_t* __restrict__ ua, const uint32_t* __restrict__ ub,
                                                   int n, const int2* __restrict__ edges,
                                                   const float4* __restrict__ x, const float4* __restrict__ p,
                                                   float th, int* __restrict__ flag, float* __restrict__ ct,
                                                   float* __restrict__ cn, double* __restrict__ cw) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int2 ea = edges[ua[i]], eb = edges[ub[i]];
    float t, w[4];
    V3 nr;
    bool hit = ee_ccd(v3(x[ea.x]), v3(x[ea.y]), v3(x[eb.x]), v3(x[eb.y]), v3(p[ea.x]), v3(p[ea.y]), v3(p[eb.x]),
                      v3(p[eb.y]), th, &t, &nr, w);
    flag[i] = hit ? 1 : 0;
    if (hit) {
        ct[i] = t;
        cn[3 * i] = nr.x;
        cn[3 * i + 1] = nr.y;
        cn[3 * i + 2] = nr.z;
        for (int k = 0; k < 4; k++) cw[4 * i + k] = w[k];
    }
}

__global__ __launch_bounds__(256) void k_compact_ee(int n, const int* __restrict__ flag, const int* __restrict__ pos,
                                                    const uint32_t* __restrict__ ua, const uint32_t* __restrict__ ub,
                                                    const int2* __restrict__ edges, const float* __restrict__ ct,
                                                    const float* __restrict__ cn, const double* __restrict__ cw,
                                                    int4* __restrict__ oq, float* __restrict__ ot,
                                                    float* __restrict__ on, double* __restrict__ ow) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !flag[i]) return;
    int o = pos[i];
    int2 ea = edges[ua[i]], eb = edges[ub[i]];
    oq[o] = make_int4(ea.x, ea.y, eb.x, eb.y);
    ot[o] = ct[i];
    for (int k = 0; k < 3; k++) on[3 * o + k] = cn[3 * i + k];
    for (int k = 0; k < 4; k++) ow[4 * o + k] = cw[4 * i + k];
}

// Edge-edge response on the current positions, for separable edges
// (ee_separable): the closest points of the two edges are pushed apart along
// the edge normal (oriented as the closest points' offset at the start of
// the step) until they are 2 x thickness apart, half each, spread over each
// edge's ends by the closest-point weights (w_k / sum w^2, so the closest
// point moves by the full half); a static edge leaves the whole gap to the
// other.  Fixes are summed in the vertex-face response's fixed-point
// accumulators.
__device__ bool ee_fix(const V3* X, const V3* q, float th, V3* d) {
    V3 side, nn;
    if (!ee_separable(X[0], X[1], X[2], X[3], q[0], q[1], q[2], q[3], th, &side, &nn)) return false;
    if (dot(side, nn) < 0.0f) nn = V3{-nn.x, -nn.y, -nn.z};
    float s, u;
    seg_closest(q[0], q[1], q[2], q[3], &s, &u);
    float gap = dot(sub(seg_point(q[0], q[1], s), seg_point(q[2], q[3], u)), nn);
    if (!(gap - 2.0f * th < -1e-6f)) return false;
    float depth = (2.0f * th - gap) * 0.5f;
    float w[4] = {1.0f - s, s, 1.0f - u, u};
    float sa = w[0] * w[0] + w[1] * w[1], sb = w[2] * w[2] + w[3] * w[3];
    for (int k = 0; k < 4; k++) {
        float f = k < 2 ? (depth * w[k]) / sa : -((depth * w[k]) / sb);
        d[k] = V3{nn.x * f, nn.y * f, nn.z * f};
    }
    return true;
}

__global__ __launch_bounds__(256) void k_resolve_ee_acc(const int4* __restrict__ cq, int nc,
                                                        const float4* __restrict__ x, const float4* __restrict__ p,
                                                        float th, unsigned long long* __restrict__ acc,
                                                        int* __restrict__ cnt) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    int4 c = cq[i];
    int id[4] = {c.x, c.y, c.z, c.w};
    float4 P[4];
    V3 X[4], q[4], d[4];
    for (int k = 0; k < 4; k++) {
        P[k] = p[id[k]];
        X[k] = v3(x[id[k]]);
        q[k] = v3(P[k]);
    }
    if (!ee_fix(X, q, th, d)) return;
    bool a_static = P[0].w == 0.0f && P[1].w == 0.0f, b_static = P[2].w == 0.0f && P[3].w == 0.0f;
    for (int k = 0; k < 4; k++) {
        if (P[k].w == 0.0f) continue;
        float sc = (k < 2 ? b_static : a_static) ? 2.0f : 1.0f;
        atomicAdd(&acc[3 * (size_t)id[k]], (unsigned long long)to_fixed(d[k].x * sc));
        atomicAdd(&acc[3 * (size_t)id[k] + 1], (unsigned long long)to_fixed(d[k].y * sc));
        atomicAdd(&acc[3 * (size_t)id[k] + 2], (unsigned long long)to_fixed(d[k].z * sc));
        atomicAdd(&cnt[id[k]], 1);
    }
}

// ---------------------------------------------------------------------------
// response
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_resolve_acc(const int* __restrict__ cvv, const int* __restrict__ cff, int nc,
                                                     const int* __restrict__ tris, const float4* __restrict__ x,
                                                     const float4* __restrict__ p, float th, float min_n2,
                                                     unsigned long long* __restrict__ acc, int* __restrict__ cnt) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    int f = cff[i];
    int id[4] = {cvv[i], tris[3 * f], tris[3 * f + 1], tris[3 * f + 2]};
    float4 P[4];
    V3 X[4], q[4];
    for (int k = 0; k < 4; k++) {
        P[k] = p[id[k]];
        X[k] = v3(x[id[k]]);
        q[k] = v3(P[k]);
    }
    // VFResolveTest (CCD.cpp:1996-2032)
    V3 n0 = vnormalize(cross(sub(X[2], X[1]), sub(X[3], X[1])));
    bool dir = dot(sub(X[0], X[1]), n0) >= 0;
    V3 cr = cross(sub(q[2], q[1]), sub(q[3], q[1]));
    if (dot(cr, cr) <= min_n2) return;
    V3 pn = vnormalize(cr);
    float s = dir ? 1.0f : -1.0f;
    V3 sn = V3{pn.x * s, pn.y * s, pn.z * s};
    float depth0 = dot(sub(q[0], q[1]), sn) - 2.0 * th;
    if (!(depth0 < -1e-6)) return;
    // VFResolveNew (CCD.cpp:2034-2090)
    V3 proj = V3{q[1].x + ((pn.x * s) * 2.0f) * th, q[1].y + ((pn.y * s) * 2.0f) * th, q[1].z + ((pn.z * s) * 2.0f) * th};
    float depth = dot(sub(proj, q[0]), sn) * 0.5;
    V3 wg = bary(q[0], q[1], q[2], q[3]);
    float sw = wg.x * wg.x + wg.y * wg.y + wg.z * wg.z;
    float wf[3] = {wg.x, wg.y, wg.z};
    float snv[3] = {sn.x, sn.y, sn.z}, pnv[3] = {pn.x, pn.y, pn.z};
    float qa[4][3] = {{q[0].x, q[0].y, q[0].z}, {q[1].x, q[1].y, q[1].z}, {q[2].x, q[2].y, q[2].z}, {q[3].x, q[3].y, q[3].z}};
    float qb[4][3];
    for (int a = 0; a < 3; a++) {
        float vfix = depth * snv[a];
        float ffix = depth * (pnv[a] * -s);
        qb[0][a] = qa[0][a] + vfix;
        qb[1][a] = qa[1][a] + (ffix * wf[0]) / sw;
        qb[2][a] = qa[2][a] + (ffix * wf[1]) / sw;
        qb[3][a] = qa[3][a] + (ffix * wf[2]) / sw;
    }
    bool face_static = (P[1].w == 0.0f) && (P[2].w == 0.0f) && (P[3].w == 0.0f);
    for (int k = 0; k < 4; k++) {
        if (P[k].w == 0.0f) continue;
        float sc = 1.0f;
        if (k == 0 && face_static) sc = 2.0f;
        if (k > 0 && P[0].w == 0.0f) sc = 2.0f;
        for (int a = 0; a < 3; a++)
            atomicAdd(&acc[3 * (size_t)id[k] + a], (unsigned long long)to_fixed((qb[k][a] - qa[k][a]) * sc));
        atomicAdd(&cnt[id[k]], 1);
    }
}

__global__ __launch_bounds__(256) void k_resolve_apply(const unsigned long long* __restrict__ acc,
                                                       const int* __restrict__ cnt, float4* __restrict__ p, int nv) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv || !cnt[v]) return;
    float4 q = p[v];
    double c = (double)cnt[v];
    q.x = q.x + (float)((double)(long long)acc[3 * (size_t)v] / 4294967296.0 / c);
    q.y = q.y + (float)((double)(long long)acc[3 * (size_t)v + 1] / 4294967296.0 / c);
    q.z = q.z + (float)((double)(long long)acc[3 * (size_t)v + 2] / 4294967296.0 / c);
    p[v] = q;
}

// ---------------------------------------------------------------------------
// host drivers
// ---------------------------------------------------------------------------
static inline int blocks(long long n) { return (int)((n + 255) / 256); }

// Working buffers only grow: detections of varying size reuse them instead
// of freeing and allocating (hipFree waits for the device) every time.
template <typename T>
static void grow(DBuf<T>& b, size_t n) {
    if (b.n < n || !b.ptr) b.alloc(std::max<size_t>(n, 1));
}

void Broadphase::build(hipStream_t s, const HashCfg& g, const float4* p, const int* tris, int nt) {
    cfg = g;
    this->nt = nt;
    grow(keys, std::max(1, nt));
    grow(ids, std::max(1, nt));
    grow(raw, std::max(1, nt));
    grow(tcell, std::max(1, nt));
    grow(k2, std::max(1, nt));
    grow(v2, std::max(1, nt));
    grow(hist, radix_hist_ints(nt));
    grow(cstart, std::max(1, g.ncells));
    grow(cend, std::max(1, g.ncells));
    if (nt > 0) hipLaunchKernelGGL(k_tri_keys, dim3(blocks(nt)), dim3(256), 0, s, g, p, tris, nt, keys.ptr, ids.ptr, raw.ptr,
                                  tcell.ptr);
    int bits = 1;
    while (bits < 32 && ((1ll << bits) < (long long)g.ncells + 1)) bits++;
    radix_sort_pairs(s, keys.ptr, ids.ptr, k2.ptr, v2.ptr, nt, bits, hist.ptr);
    hipLaunchKernelGGL(k_fill_i32, dim3(blocks(g.ncells)), dim3(256), 0, s, cstart.ptr, g.ncells, -1);
    hipLaunchKernelGGL(k_fill_i32, dim3(blocks(g.ncells)), dim3(256), 0, s, cend.ptr, g.ncells, -1);
    if (nt > 0) hipLaunchKernelGGL(k_cell_ranges, dim3(blocks(nt)), dim3(256), 0, s, keys.ptr, nt, cstart.ptr, cend.ptr);
    PBD_HIP(hipGetLastError());
}

long long Broadphase::pairs(hipStream_t s, const float4* p, const int* tris, int nv) {
    npairs = 0;
    if (nt <= 0) return 0;
    grow(cnt, nt + 1);
    grow(off, nt + 1);
    grow(scan_tmp, (nt + SB) / SB + 2);
    hipLaunchKernelGGL(k_pairs<0>, dim3(blocks(nt)), dim3(256), 0, s, cfg, p, tris, nt, raw.ptr, ids.ptr, cstart.ptr,
                       cend.ptr, tcell.ptr, cnt.ptr, (uint32_t*)nullptr, (uint32_t*)nullptr);
    PBD_HIP(hipMemsetAsync(cnt.ptr + nt, 0, sizeof(int), s));
    device_scan(s, cnt.ptr, off.ptr, nt + 1, scan_tmp.ptr);
    int total = 0;
    PBD_HIP(hipMemcpyAsync(&total, off.ptr + nt, sizeof(int), hipMemcpyDeviceToHost, s));
    PBD_HIP(hipStreamSynchronize(s));
    PBD_CHECK(total >= 0, PBD_ERR_LIMIT, "candidate pair count overflows int32");
    st_dets++;
    st_tris += nt;
    st_entries += nt;
    st_raw += total;
    if (total == 0) return 0;
    grow(pv, total);
    grow(pf, total);
    grow(t1, total);
    grow(t2, total);
    grow(flag, total + 1);
    grow(pos, total + 1);
    hipLaunchKernelGGL(k_pairs<1>, dim3(blocks(nt)), dim3(256), 0, s, cfg, p, tris, nt, raw.ptr, ids.ptr, cstart.ptr,
                       cend.ptr, tcell.ptr, off.ptr, pv.ptr, pf.ptr);
    return unique_pairs(s, total, nt, nv);
}

// Sorts the `total` raw pairs in pv/pf by (v, f) and keeps one of each in uv/uf.
long long Broadphase::unique_pairs(hipStream_t s, int total, int nt, int nv) {
    long long uniq = unique_into(s, total, nt, nv, uv, uf);
    npairs = uniq;
    st_pairs += uniq;
    return uniq;
}

// Sorts the `total` raw pairs in pv/pf by (pv, pf) (pv < nv, pf < nt) and
// writes one of each to oa/ob.
long long Broadphase::unique_into(hipStream_t s, int total, int nt, int nv, DBuf<uint32_t>& oa, DBuf<uint32_t>& ob) {
    // sort by (v, f): by f first, then stably by v
    int fb = 1, vb = 1;
    while (fb < 31 && (1ll << fb) < nt) fb++;
    while (vb < 31 && (1ll << vb) < nv) vb++;
    grow(hist2, radix_hist_ints(total));
    radix_sort_pairs(s, pf.ptr, pv.ptr, t1.ptr, t2.ptr, total, fb, hist2.ptr);
    radix_sort_pairs(s, pv.ptr, pf.ptr, t1.ptr, t2.ptr, total, vb, hist2.ptr);
    hipLaunchKernelGGL(k_unique_flags, dim3(blocks(total)), dim3(256), 0, s, pv.ptr, pf.ptr, total, flag.ptr);
    PBD_HIP(hipMemsetAsync(flag.ptr + total, 0, sizeof(int), s));
    grow(scan_tmp, (total + 1 + SB) / SB + 2);
    device_scan(s, flag.ptr, pos.ptr, total + 1, scan_tmp.ptr);
    int uniq = 0;
    PBD_HIP(hipMemcpyAsync(&uniq, pos.ptr + total, sizeof(int), hipMemcpyDeviceToHost, s));
    PBD_HIP(hipStreamSynchronize(s));
    grow(oa, std::max(1, uniq));
    grow(ob, std::max(1, uniq));
    hipLaunchKernelGGL(k_compact_pairs, dim3(blocks(total)), dim3(256), 0, s, pv.ptr, pf.ptr, total, flag.ptr, pos.ptr,
                       oa.ptr, ob.ptr);
    PBD_HIP(hipGetLastError());
    return uniq;
}

void Broadphase::set_edges(hipStream_t s, const std::vector<int>& te, const std::vector<int2>& ed) {
    tedge.upload(te, s);
    edges.upload(ed, s);
    nedges = (int)ed.size();
    // vertex -> neighbour CSR from the edges (each edge once, both directions)
    int nv = 0;
    for (const int2& e : ed) nv = std::max(nv, e.y + 1);
    std::vector<int> off((size_t)nv + 1, 0), nbr((size_t)2 * ed.size());
    for (const int2& e : ed) {
        off[e.x + 1]++;
        off[e.y + 1]++;
    }
    for (int v = 0; v < nv; v++) off[v + 1] += off[v];
    std::vector<int> fill(off.begin(), off.end() - 1);
    for (const int2& e : ed) {
        nbr[fill[e.x]++] = e.y;
        nbr[fill[e.y]++] = e.x;
    }
    voff.upload(off, s);
    vnbr.upload(nbr, s);
    PBD_HIP(hipStreamSynchronize(s));
}

// Edge-edge candidates from the cell table of the last build() (see k_ee_pairs).
long long Broadphase::ee_pairs(hipStream_t s, const float4* p, const int* tris) {
    ee_npairs = 0;
    ee_ncontacts = 0;
    if (nt <= 0) return 0;
    PBD_CHECK(tedge.n >= (size_t)3 * nt, PBD_ERR_STATE, "internal: edge table does not match the triangles");
    grow(cnt, nt + 1);
    grow(off, nt + 1);
    grow(scan_tmp, (nt + SB) / SB + 2);
    hipLaunchKernelGGL(k_ee_pairs<0>, dim3(blocks(nt)), dim3(256), 0, s, cfg, p, tris, nt, raw.ptr, ids.ptr,
                       cstart.ptr, cend.ptr, tcell.ptr, tedge.ptr, edges.ptr, voff.ptr, vnbr.ptr, cnt.ptr,
                       (uint32_t*)nullptr, (uint32_t*)nullptr);
    PBD_HIP(hipMemsetAsync(cnt.ptr + nt, 0, sizeof(int), s));
    device_scan(s, cnt.ptr, off.ptr, nt + 1, scan_tmp.ptr);
    int total = 0;
    PBD_HIP(hipMemcpyAsync(&total, off.ptr + nt, sizeof(int), hipMemcpyDeviceToHost, s));
    PBD_HIP(hipStreamSynchronize(s));
    PBD_CHECK(total >= 0, PBD_ERR_LIMIT, "edge-edge candidate count overflows int32");
    st_ee_raw += total;
    if (total == 0) return 0;
    grow(pv, total);
    grow(pf, total);
    grow(t1, total);
    grow(t2, total);
    grow(flag, total + 1);
    grow(pos, total + 1);
    hipLaunchKernelGGL(k_ee_pairs<1>, dim3(blocks(nt)), dim3(256), 0, s, cfg, p, tris, nt, raw.ptr, ids.ptr,
                       cstart.ptr, cend.ptr, tcell.ptr, tedge.ptr, edges.ptr, voff.ptr, vnbr.ptr, off.ptr, pv.ptr,
                       pf.ptr);
    ee_npairs = unique_into(s, total, nedges, nedges, eua, eub);
    st_ee_pairs += ee_npairs;
    return ee_npairs;
}

long long Broadphase::ee_narrow(hipStream_t s, const float4* x, const float4* p, float th) {
    ee_ncontacts = 0;
    if (ee_npairs <= 0) return 0;
    long long n = ee_npairs;
    grow(flag, n + 1);
    grow(pos, n + 1);
    grow(ctmp_t, n);
    grow(ctmp_n, 3 * n);
    grow(ctmp_w, 4 * n);
    hipLaunchKernelGGL(k_narrow_ee, dim3(blocks(n)), dim3(256), 0, s, eua.ptr, eub.ptr, (int)n, edges.ptr, x, p, th,
                       flag.ptr, ctmp_t.ptr, ctmp_n.ptr, ctmp_w.ptr);
    PBD_HIP(hipMemsetAsync(flag.ptr + n, 0, sizeof(int), s));
    grow(scan_tmp, (n + 1 + SB) / SB + 2);
    device_scan(s, flag.ptr, pos.ptr, (int)n + 1, scan_tmp.ptr);
    int nc = 0;
    PBD_HIP(hipMemcpyAsync(&nc, pos.ptr + n, sizeof(int), hipMemcpyDeviceToHost, s));
    PBD_HIP(hipStreamSynchronize(s));
    grow(ecq, std::max(1, nc));
    grow(ect, std::max(1, nc));
    grow(ecn, 3 * std::max(1, nc));
    grow(ecw, 4 * std::max(1, nc));
    hipLaunchKernelGGL(k_compact_ee, dim3(blocks(n)), dim3(256), 0, s, (int)n, flag.ptr, pos.ptr, eua.ptr, eub.ptr,
                       edges.ptr, ctmp_t.ptr, ctmp_n.ptr, ctmp_w.ptr, ecq.ptr, ect.ptr, ecn.ptr, ecw.ptr);
    PBD_HIP(hipGetLastError());
    ee_ncontacts = nc;
    st_ee_contacts += nc;
    return nc;
}

long long Broadphase::pairs_swept(hipStream_t s, const HashCfg& g, const float4* x, const float4* p, const int* tris,
                                  int nt_, int nv, float th) {
    cfg = g;
    nt = nt_;
    npairs = 0;
    if (nt <= 0 || nv <= 0) return 0;
    grow(sw_box, 2 * (size_t)nt);
    grow(sw_cell, 2 * (size_t)nt);
    grow(cnt, (size_t)std::max(nt, nv) + 1);
    grow(off, (size_t)std::max(nt, nv) + 1);
    grow(sw_err, 2);
    grow(sw_vin, nv);
    grow(scan_tmp, (std::max(nt, nv) + SB) / SB + 2);
    PBD_HIP(hipMemsetAsync(sw_err.ptr, 0, 2 * sizeof(int), s));
    hipLaunchKernelGGL(k_swept_faces, dim3(blocks(nt)), dim3(256), 0, s, g, x, p, tris, nt, 4.0f * th, sw_box.ptr,
                       sw_cell.ptr, cnt.ptr, sw_err.ptr);
    PBD_HIP(hipMemsetAsync(cnt.ptr + nt, 0, sizeof(int), s));
    device_scan(s, cnt.ptr, off.ptr, nt + 1, scan_tmp.ptr);
    int hst[3] = {0, 0, 0};
    PBD_HIP(hipMemcpyAsync(&hst[0], off.ptr + nt, sizeof(int), hipMemcpyDeviceToHost, s));
    PBD_HIP(hipMemcpyAsync(&hst[1], sw_err.ptr, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
    PBD_HIP(hipStreamSynchronize(s));
    PBD_CHECK(!hst[1], PBD_ERR_LIMIT, "a face's swept box spans more than 4096 cells: use a larger cell");
    const int ne = hst[0];
    PBD_CHECK(ne >= 0, PBD_ERR_LIMIT, "swept-box cell entries overflow int32");
    grow(keys, ne);
    grow(ids, ne);
    grow(k2, ne);
    grow(v2, ne);
    grow(hist, radix_hist_ints(ne));
    grow(cstart, std::max(1, g.ncells));
    grow(cend, std::max(1, g.ncells));
    hipLaunchKernelGGL(k_swept_entries, dim3(blocks(nt)), dim3(256), 0, s, g, nt, sw_cell.ptr, off.ptr, keys.ptr, ids.ptr);
    int bits = 1;
    while (bits < 32 && ((1ll << bits) < (long long)g.ncells + 1)) bits++;
    radix_sort_pairs(s, keys.ptr, ids.ptr, k2.ptr, v2.ptr, ne, bits, hist.ptr);
    hipLaunchKernelGGL(k_fill_i32, dim3(blocks(g.ncells)), dim3(256), 0, s, cstart.ptr, g.ncells, -1);
    hipLaunchKernelGGL(k_fill_i32, dim3(blocks(g.ncells)), dim3(256), 0, s, cend.ptr, g.ncells, -1);
    if (ne > 0) hipLaunchKernelGGL(k_cell_ranges, dim3(blocks(ne)), dim3(256), 0, s, keys.ptr, ne, cstart.ptr, cend.ptr);
    PBD_HIP(hipMemsetAsync(sw_vin.ptr, 0, sizeof(int) * (size_t)nv, s));
    hipLaunchKernelGGL(k_mark_vertices, dim3(blocks(3ll * nt)), dim3(256), 0, s, tris, nt, sw_vin.ptr);
    hipLaunchKernelGGL(k_swept_pairs<0>, dim3(blocks(nv)), dim3(256), 0, s, g, x, p, sw_vin.ptr, nv, tris, sw_box.ptr,
                       sw_cell.ptr, ids.ptr, cstart.ptr, cend.ptr, cnt.ptr, (uint32_t*)nullptr, (uint32_t*)nullptr,
                       sw_err.ptr);
    PBD_HIP(hipMemsetAsync(cnt.ptr + nv, 0, sizeof(int), s));
    device_scan(s, cnt.ptr, off.ptr, nv + 1, scan_tmp.ptr);
    PBD_HIP(hipMemcpyAsync(&hst[0], off.ptr + nv, sizeof(int), hipMemcpyDeviceToHost, s));
    PBD_HIP(hipMemcpyAsync(&hst[1], sw_err.ptr, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
    PBD_HIP(hipStreamSynchronize(s));
    PBD_CHECK(!hst[2], PBD_ERR_LIMIT, "a vertex's swept box spans more than 4096 cells: use a larger cell");
    const int total = hst[0];
    PBD_CHECK(total >= 0, PBD_ERR_LIMIT, "candidate pair count overflows int32");
    st_dets++;
    st_tris += nt;
    st_entries += ne;
    st_raw += total;
    if (total == 0) return 0;
    grow(pv, total);
    grow(pf, total);
    grow(t1, total);
    grow(t2, total);
    grow(flag, total + 1);
    grow(pos, total + 1);
    hipLaunchKernelGGL(k_swept_pairs<1>, dim3(blocks(nv)), dim3(256), 0, s, g, x, p, sw_vin.ptr, nv, tris, sw_box.ptr,
                       sw_cell.ptr, ids.ptr, cstart.ptr, cend.ptr, off.ptr, pv.ptr, pf.ptr, sw_err.ptr);
    return unique_pairs(s, total, nt, nv);
}

long long Broadphase::narrow(hipStream_t s, const float4* x, const float4* p, const int* tris, float th) {
    ncontacts = 0;
    if (npairs <= 0) return 0;
    long long n = npairs;
    grow(flag, n + 1);
    grow(pos, n + 1);
    grow(ctmp_t, n);
    grow(ctmp_n, 3 * n);
    grow(ctmp_w, 4 * n);
    hipLaunchKernelGGL(k_narrow, dim3(blocks(n)), dim3(256), 0, s, uv.ptr, uf.ptr, (int)n, tris, x, p, th, flag.ptr,
                       ctmp_t.ptr, ctmp_n.ptr, ctmp_w.ptr);
    PBD_HIP(hipMemsetAsync(flag.ptr + n, 0, sizeof(int), s));
    grow(scan_tmp, (n + 1 + SB) / SB + 2);
    device_scan(s, flag.ptr, pos.ptr, (int)n + 1, scan_tmp.ptr);
    int nc = 0;
    PBD_HIP(hipMemcpyAsync(&nc, pos.ptr + n, sizeof(int), hipMemcpyDeviceToHost, s));
    PBD_HIP(hipStreamSynchronize(s));
    grow(cv, std::max(1, nc));
    grow(cf, std::max(1, nc));
    grow(ct, std::max(1, nc));
    grow(cn, 3 * std::max(1, nc));
    grow(cw, 4 * std::max(1, nc));
    hipLaunchKernelGGL(k_compact_contacts, dim3(blocks(n)), dim3(256), 0, s, (int)n, flag.ptr, pos.ptr, uv.ptr, uf.ptr,
                       ctmp_t.ptr, ctmp_n.ptr, ctmp_w.ptr, cv.ptr, cf.ptr, ct.ptr, cn.ptr, cw.ptr);
    PBD_HIP(hipGetLastError());
    ncontacts = nc;
    st_contacts += nc;
    return nc;
}

void Broadphase::resolve(hipStream_t s, const float4* x, float4* p, const int* tris, int nv, float th, int passes,
                         float min_n2) {
    if (ncontacts <= 0 && ee_ncontacts <= 0) return;
    st_resolved += (ncontacts + ee_ncontacts) * passes;
    grow(acc, (size_t)3 * nv);
    grow(vcnt, nv);
    for (int ps = 0; ps < passes; ps++) {
        PBD_HIP(hipMemsetAsync(acc.ptr, 0, sizeof(unsigned long long) * 3 * (size_t)nv, s));
        PBD_HIP(hipMemsetAsync(vcnt.ptr, 0, sizeof(int) * (size_t)nv, s));
        if (ncontacts > 0)
            hipLaunchKernelGGL(k_resolve_acc, dim3(blocks(ncontacts)), dim3(256), 0, s, cv.ptr, cf.ptr, (int)ncontacts,
                               tris, x, (const float4*)p, th, min_n2, acc.ptr, vcnt.ptr);
        if (ee_ncontacts > 0)
            hipLaunchKernelGGL(k_resolve_ee_acc, dim3(blocks(ee_ncontacts)), dim3(256), 0, s, ecq.ptr,
                               (int)ee_ncontacts, x, (const float4*)p, th, acc.ptr, vcnt.ptr);
        hipLaunchKernelGGL(k_resolve_apply, dim3(blocks(nv)), dim3(256), 0, s, acc.ptr, vcnt.ptr, p, nv);
    }
    PBD_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Broadphase build without allocation (pbd_sh_build_noalloc).  The caller
// brings device positions (3 floats per vertex), triangles and the output
// arrays; the radix sort runs in the caller's key and id arrays, with its
// ping-pong copies and digit counts in `scratch`, one static array per GPU
// that the process keeps until it ends.  Every call overwrites what it uses
// (digit counts and scans are written in full before they are read), so
// nothing carries over between calls; calls that overlap on one GPU race on
// it.
// ---------------------------------------------------------------------------
constexpr size_t SH_NOALLOC_HIST_INTS = 2 * ((size_t)256 * (SH_NOALLOC_MAX_TRIS / 4096)) +
                                        ((size_t)256 * (SH_NOALLOC_MAX_TRIS / 4096) + 1023) / 1024 + 2;
constexpr size_t SH_NOALLOC_WORDS = 2 * (size_t)SH_NOALLOC_MAX_TRIS + SH_NOALLOC_HIST_INTS;
static __device__ __align__(256) uint32_t scratch[SH_NOALLOC_WORDS];

size_t sh_noalloc_scratch_bytes() { return sizeof(uint32_t) * SH_NOALLOC_WORDS; }

// Device address of this GPU's `scratch`, resolved once per device with the
// device current.  The host-side symbol is not a device address: only the
// resolved pointer is ever passed to a kernel.
static void* scratch_address(int device) {
    static std::mutex mu;
    static std::vector<void*> addr;
    std::lock_guard<std::mutex> lock(mu);
    if ((int)addr.size() <= device) addr.resize((size_t)device + 1, nullptr);
    if (!addr[device]) {
        void* p = nullptr;
        PBD_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        PBD_CHECK(p != nullptr, PBD_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
        addr[device] = p;
    }
    return addr[device];
}

// k_tri_keys on 3-float positions; a triangle with a vertex index outside
// [0, nv) is given key -1 (outside the grid) instead of being read.
__global__ __launch_bounds__(256) void k_tri_keys_noalloc(HashCfg g, const float* __restrict__ p, int nv,
                                                          const int* __restrict__ tris, int nt,
                                                          uint32_t* __restrict__ keys1, uint32_t* __restrict__ ids,
                                                          int* __restrict__ raw) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    const int a = tris[3 * t], b = tris[3 * t + 1], c = tris[3 * t + 2];
    int k = -1;
    if (a >= 0 && a < nv && b >= 0 && b < nv && c >= 0 && c < nv) {
        int gp[3];
        tri_cell(g, p, tris, t, gp);
        k = cell_key(g, gp[0], gp[1], gp[2]);
    }
    keys1[t] = (uint32_t)(k + 1);
    ids[t] = (uint32_t)t;
    raw[t] = k;
}

// keys1 (key + 1, the sort's unsigned form) -> key, in place.
__global__ __launch_bounds__(256) void k_keys_signed(uint32_t* __restrict__ keys1, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) keys1[i] = (uint32_t)((int)keys1[i] - 1);
}

void sh_build_noalloc(hipStream_t s, int device, const HashCfg& g, const float* pos, int nv, const int* tris, int nt,
                      int* sorted_keys, int* sorted_ids, int* raw_keys, int* cell_start, int* cell_end) {
    PBD_CHECK(nt <= SH_NOALLOC_MAX_TRIS, PBD_ERR_LIMIT,
              "pbd_sh_build_noalloc: " + std::to_string(nt) + " triangles exceed its maximum of " +
                  std::to_string(SH_NOALLOC_MAX_TRIS));
    uint32_t* scr = (uint32_t*)scratch_address(device);
    uint32_t* k2 = scr;
    uint32_t* v2 = scr + SH_NOALLOC_MAX_TRIS;
    int* hist = (int*)(scr + 2 * (size_t)SH_NOALLOC_MAX_TRIS);
    uint32_t* k = (uint32_t*)sorted_keys;
    uint32_t* v = (uint32_t*)sorted_ids;
    if (nt > 0)
        hipLaunchKernelGGL(k_tri_keys_noalloc, dim3(blocks(nt)), dim3(256), 0, s, g, pos, nv, tris, nt, k, v, raw_keys);
    int bits = 1;
    while (bits < 32 && ((1ll << bits) < (long long)g.ncells + 1)) bits++;
    radix_sort_pairs(s, k, v, k2, v2, nt, bits, hist);
    hipLaunchKernelGGL(k_fill_i32, dim3(blocks(g.ncells)), dim3(256), 0, s, cell_start, g.ncells, -1);
    hipLaunchKernelGGL(k_fill_i32, dim3(blocks(g.ncells)), dim3(256), 0, s, cell_end, g.ncells, -1);
    if (nt > 0) {
        hipLaunchKernelGGL(k_cell_ranges, dim3(blocks(nt)), dim3(256), 0, s, k, nt, cell_start, cell_end);
        hipLaunchKernelGGL(k_keys_signed, dim3(blocks(nt)), dim3(256), 0, s, k, nt);
    }
    PBD_HIP(hipGetLastError());
}

}  // namespace pbd
