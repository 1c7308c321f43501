// Substep kernels for gfx950: advect, colour-batched Gauss-Seidel projection
// (global per-colour launches and LDS tiles), analytic colliders, FP64
// residual reduction, integrate.
//
// Vertex state is float4 (x, y, z, w) with w the weight the reference keeps in
// massBuffer (1 free, 0 pinned), so one 16-byte load brings a position and
// its weight.  Reference paths are relative to
// PositionBasedDynamicsCUDA/PositionBasedDynamicsCUDA/.
#include "pbd_internal.hpp"
#include "pbd_math.hpp"

namespace pbd {

__global__ __launch_bounds__(256) void k_advect(const ClothDev* __restrict__ cl,
                                                const int* __restrict__ vcloth,
                                                const float4* __restrict__ x, float4* __restrict__ v,
                                                float4* __restrict__ p, int b, int e, float dt,
                                                int skip_whole) {
    int i = b + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= e) return;
    const ClothDev& c = cl[vcloth[i]];
    if (skip_whole && c.whole) return;
    float4 V;
    p[i] = advect_one(c, x[i], v[i], dt, &V);
    v[i] = V;
}

void launch_advect(hipStream_t s, const ClothDev* cloths, const int* vcloth, const float4* x,
                   float4* v, float4* p, int b, int e, float dt, int skip_whole) {
    if (e <= b) return;
    hipLaunchKernelGGL(k_advect, dim3((e - b + 255) / 256), dim3(256), 0, s, cloths, vcloth, x, v, p, b, e, dt,
                       skip_whole);
    PBD_HIP(hipGetLastError());
}

// IntegrationGPUKernel (PBD_cuda.cu:1467-1482): v = (p - x)/dt ; x = p.
// p of vertex i is read from buffer pfinal[cloth] (pend when pfinal is null)
// and, when that is not pend, also stored into pend so that the current p
// buffer holds every cloth's result.  skip_whole: single-tile cloths were
// integrated by the tile kernel already (only their p is carried over).
__global__ __launch_bounds__(256) void k_integrate(const ClothDev* __restrict__ cl,
                                                   const int* __restrict__ vcloth, float4* __restrict__ x,
                                                   float4* __restrict__ v, float4* __restrict__ p0,
                                                   float4* __restrict__ p1, const int* __restrict__ pfinal, int pend,
                                                   int b, int e, float dt, int skip_whole) {
    int i = b + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= e) return;
    const int c = vcloth[i];
    const int src = pfinal ? pfinal[c] : pend;
    const float4 P = (src ? p1 : p0)[i];
    if (!(skip_whole && cl[c].whole)) {
        const float4 X = x[i];
        v[i] = make_float4((P.x - X.x) / dt, (P.y - X.y) / dt, (P.z - X.z) / dt, 0.0f);
        x[i] = make_float4(P.x, P.y, P.z, X.w);
    }
    if (src != pend) (pend ? p1 : p0)[i] = P;
}

void launch_integrate(hipStream_t s, const ClothDev* cloths, const int* vcloth, float4* x, float4* v, float4* p0,
                      float4* p1, const int* pfinal, int pend, int b, int e, float dt, int skip_whole) {
    if (e <= b) return;
    hipLaunchKernelGGL(k_integrate, dim3((e - b + 255) / 256), dim3(256), 0, s, cloths, vcloth, x, v, p0, p1, pfinal,
                       pend, b, e, dt, skip_whole);
    PBD_HIP(hipGetLastError());
}

// Copies each cloth's projected p from buffer pfinal[cloth] into pend.
__global__ __launch_bounds__(256) void k_consolidate(const int* __restrict__ vcloth, float4* __restrict__ p0,
                                                     float4* __restrict__ p1, const int* __restrict__ pfinal,
                                                     int pend, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int src = pfinal[vcloth[i]];
    if (src != pend) (pend ? p1 : p0)[i] = (src ? p1 : p0)[i];
}

void launch_consolidate(hipStream_t s, const int* vcloth, float4* p0, float4* p1, const int* pfinal, int pend,
                        int n) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_consolidate, dim3((n + 255) / 256), dim3(256), 0, s, vcloth, p0, p1, pfinal, pend, n);
    PBD_HIP(hipGetLastError());
}

// Which p buffer holds each cloth after a tile projection that started in
// buffer `start`: single-tile cloths wrote start ^ 1; multi-tile cloths
// flipped once per chunk they ran -- ceil(min(cap, iterations) / ts) chunks
// when a per-cloth budget ends them early without copying, else every one
// of the `launched` chunks.
__global__ void k_final_buffers(const ClothDev* __restrict__ cl, int ncl, int start, int launched, int ts,
                                int iterations, const int* __restrict__ cap, int* __restrict__ pfinal) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncl) return;
    int run = launched;
    if (cap) {
        const int k = min(cap[c], iterations);
        run = (k + ts - 1) / ts;
    }
    pfinal[c] = cl[c].whole ? (start ^ 1) : (start ^ (run & 1));
}

void launch_final_buffers(hipStream_t s, const ClothDev* cl, int ncl, int start, int launched, int ts,
                          int iterations, const int* cap, int* pfinal) {
    if (ncl <= 0) return;
    hipLaunchKernelGGL(k_final_buffers, dim3((ncl + 255) / 256), dim3(256), 0, s, cl, ncl, start, launched, ts,
                       iterations, cap, pfinal);
    PBD_HIP(hipGetLastError());
}

// One colour of ProjectContraintsGPUKernel (PBD_cuda.cu:1240-1306) over the
// whole batch: distance records (i0, i1, rest bits, cloth) and bending records
// (idx4, (rest, cloth bits)).  Unlike the reference launch (PBD_cuda.cu:1332-1337)
// the grid covers every prim of the colour.
__global__ __launch_bounds__(256) void k_project_colour(const int4* __restrict__ dist, int db, int de,
                                                        const int4* __restrict__ bidx,
                                                        const float2* __restrict__ brc, int bb, int be,
                                                        const float* __restrict__ ktab_d,
                                                        const float* __restrict__ ktab_b, int kmax,
                                                        int sweep, const int* __restrict__ cap,
                                                        float4* __restrict__ p) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    int nd = de - db;
    int ks = sweep < kmax ? sweep : kmax - 1;
    if (j < nd) {
        int4 r = dist[db + j];
        if (cap && sweep >= cap[r.w]) return;  // this cloth's sweep budget is spent
        float k = ktab_d[r.w * kmax + ks];
        float4 a = p[r.x], b = p[r.y];
        project_distance(a, b, __int_as_float(r.z), k);
        p[r.x] = a;
        p[r.y] = b;
        return;
    }
    j -= nd;
    if (j < be - bb) {
        int4 q = bidx[bb + j];
        float2 rc = brc[bb + j];
        if (cap && sweep >= cap[__float_as_int(rc.y)]) return;
        float k = ktab_b[__float_as_int(rc.y) * kmax + ks];
        float4 q0 = p[q.x], q1 = p[q.y], q2 = p[q.z], q3 = p[q.w];
        project_bending(q0, q1, q2, q3, rc.x, k);
        p[q.x] = q0;
        p[q.y] = q1;
        p[q.z] = q2;
        p[q.w] = q3;
    }
}

void launch_project_colour(hipStream_t s, const int4* dist, int db, int de, const int4* bidx,
                           const float2* brc, int bb, int be, const float* ktab_d,
                           const float* ktab_b, int kmax, int sweep, const int* cap, float4* p) {
    int n = (de - db) + (be - bb);
    if (n <= 0) return;
    hipLaunchKernelGGL(k_project_colour, dim3((n + 255) / 256), dim3(256), 0, s, dist, db, de, bidx, brc, bb,
                       be, ktab_d, ktab_b, kmax, sweep, cap, p);
    PBD_HIP(hipGetLastError());
}

__global__ __launch_bounds__(256) void k_colliders(const ClothDev* __restrict__ cl,
                                                   const int* __restrict__ vcloth, float4* __restrict__ p,
                                                   int b, int e, const int* __restrict__ cap, int sweep) {
    int i = b + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= e) return;
    const int ci = vcloth[i];
    if (cap && sweep >= cap[ci]) return;
    const ClothDev& c = cl[ci];
    if (!c.sphere_on && !c.box_on) return;
    float4 q = p[i];
    collide_point(q, c.sphere_on, c.sphere_c, c.sphere_r, c.box_on, c.box_lo, c.box_hi);
    p[i] = q;
}

void launch_colliders(hipStream_t s, const ClothDev* cloths, const int* vcloth, float4* p, int b, int e,
                      const int* cap, int sweep) {
    if (e <= b) return;
    hipLaunchKernelGGL(k_colliders, dim3((e - b + 255) / 256), dim3(256), 0, s, cloths, vcloth, p, b, e, cap,
                       sweep);
    PBD_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Jacobi sweep with constraint averaging (the reference declares
// SolverType::JACOBI, PBD_Basic.cuh:46-51, and never implements it; this is
// the averaged Jacobi of Macklin et al. 2014, "Unified particle physics for
// real-time applications", sec. 4.2).  One thread per vertex gathers, from
// p_in, the correction each incident prim (ascending prim id: distance prims,
// then bending prims) would give it when projected alone, with the prim's
// own vertex order and the Gauss-Seidel projection arithmetic, and moves by
// omega/n times their sum (n = incident prims).  Pinned vertices (w = 0) and
// cloths whose sweep budget is spent are copied.  Colliders follow in the
// same thread.  The result depends on no launch or summation order beyond
// the per-vertex prim order.
//
// Distance incidences, 8 B each: (other vertex | role << 31, rest bits),
// role 1 when this vertex is the prim's second.  The first JAC_SLOTS of a
// vertex are slot-major (slot k of vertex i at dslot[k * nv + i], so a
// wave's loads are contiguous), the rest in a per-vertex overflow list
// (doff).  Bending incidences: boff / bent (prim vertices) / bmeta (rest
// bits, role).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_jacobi(JacobiDev J, const ClothDev* __restrict__ cl,
                                                const int* __restrict__ vcloth, const float* __restrict__ ktab_d,
                                                const float* __restrict__ ktab_b, int kmax, int sweep,
                                                const int* __restrict__ cap, float omega,
                                                const float4* __restrict__ pin, float4* __restrict__ pout, int b,
                                                int e) {
    int i = b + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= e) return;
    float4 pi = pin[i];
    const int ci = vcloth[i];
    if (pi.w == 0.0f || (cap && sweep >= cap[ci])) {
        pout[i] = pi;
        return;
    }
    const int ks = sweep < kmax ? sweep : kmax - 1;
    const float kd = ktab_d[ci * kmax + ks];
    const int deg = J.ddeg[i];
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    auto dist_one = [&](int2 d) {
        const int role = (int)((unsigned)d.x >> 31);
        float4 po = pin[d.x & 0x7fffffff];
        float4 a = role ? po : pi, bb = role ? pi : po;
        project_distance(a, bb, __int_as_float(d.y), kd);
        float4 r = role ? bb : a;
        sx = sx + (r.x - pi.x);
        sy = sy + (r.y - pi.y);
        sz = sz + (r.z - pi.z);
    };
    const int ns = deg < JAC_SLOTS ? deg : JAC_SLOTS;
    for (int k = 0; k < ns; k++) dist_one(J.dslot[(size_t)k * J.nv + i]);
    for (int j = J.doff[i]; j < J.doff[i + 1]; j++) dist_one(J.dovf[j]);
    const int b0 = J.boff[i], b1 = J.boff[i + 1];
    if (b1 > b0) {
        const float kb = ktab_b[ci * kmax + ks];
        for (int j = b0; j < b1; j++) {
            int4 q = J.bent[j];
            int2 m = J.bmeta[j];
            float4 q0 = pin[q.x], q1 = pin[q.y], q2 = pin[q.z], q3 = pin[q.w];
            project_bending(q0, q1, q2, q3, __int_as_float(m.x), kb);
            float4 r = m.y == 0 ? q0 : m.y == 1 ? q1 : m.y == 2 ? q2 : q3;
            sx = sx + (r.x - pi.x);
            sy = sy + (r.y - pi.y);
            sz = sz + (r.z - pi.z);
        }
    }
    const int n = deg + (b1 - b0);
    if (n > 0) {
        const float f = omega / (float)n;
        pi.x = pi.x + f * sx;
        pi.y = pi.y + f * sy;
        pi.z = pi.z + f * sz;
    }
    const ClothDev& c = cl[ci];
    if (c.sphere_on || c.box_on) collide_point(pi, c.sphere_on, c.sphere_c, c.sphere_r, c.box_on, c.box_lo, c.box_hi);
    pout[i] = pi;
}

void launch_jacobi(hipStream_t s, const JacobiDev& J, const ClothDev* cloths, const int* vcloth, const float* ktab_d,
                   const float* ktab_b, int kmax, int sweep, const int* cap, float omega, const float4* pin,
                   float4* pout, int b, int e) {
    if (e <= b) return;
    hipLaunchKernelGGL(k_jacobi, dim3((e - b + 255) / 256), dim3(256), 0, s, J, cloths, vcloth, ktab_d, ktab_b, kmax,
                       sweep, cap, omega, pin, pout, b, e);
    PBD_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// FP64 residual r = |d - L| / L per distance prim, d from the FP32 positions
// widened to double.  Max per cloth via 64-bit atomicMax on the bit pattern
// (r >= 0, so IEEE order equals unsigned order); sum of r^2 per cloth.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double rel_residual(float4 a, float4 b, float L) {
    double dx = (double)a.x - (double)b.x;
    double dy = (double)a.y - (double)b.y;
    double dz = (double)a.z - (double)b.z;
    double d = sqrt(dx * dx + dy * dy + dz * dz);
    double l = (double)L;
    return fabs(d - l) / l;
}

__device__ __forceinline__ double wave_max(double v) {
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
    return v;
}
__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

__global__ __launch_bounds__(256) void k_residual(const int4* __restrict__ dist, int n,
                                                  const float4* __restrict__ p,
                                                  unsigned long long* __restrict__ rmax,
                                                  double* __restrict__ rsum) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    int cl = -1;
    double r = 0.0;
    if (j < n) {
        int4 q = dist[j];
        cl = q.w;
        r = rel_residual(p[q.x], p[q.y], __int_as_float(q.z));
    }
    int c0 = __shfl(cl, 0);
    unsigned long long same = __ballot(cl == c0 || cl < 0);
    if (same == ~0ull) {
        double m = wave_max(r), s = wave_sum(r * r);
        if ((threadIdx.x & 63) == 0 && c0 >= 0) {
            atomicMax(&rmax[c0], (unsigned long long)__double_as_longlong(m));
            atomicAdd(&rsum[c0], s);
        }
    } else if (cl >= 0) {
        atomicMax(&rmax[cl], (unsigned long long)__double_as_longlong(r));
        atomicAdd(&rsum[cl], r * r);
    }
}

void launch_residual(hipStream_t s, const int4* dist, int n, const float4* p, unsigned long long* rmax,
                     double* rsum, int*) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_residual, dim3((n + 255) / 256), dim3(256), 0, s, dist, n, p, rmax, rsum);
    PBD_HIP(hipGetLastError());
}

// Plain streaming copy (HBM calibration): each float4 read once, written once.
// Streaming copy: four 16-byte elements per thread per step, all loads
// before the stores, non-temporal (the data is not read again soon).  On the
// 1 GiB calibration copy: 6.2 TB/s read + write, against 4.7 TB/s for one
// element per thread per step with ordinary accesses.
typedef float f32x4v __attribute__((ext_vector_type(4)));
constexpr int COPY_UNROLL = 4;

__global__ __launch_bounds__(256) void k_copy(const float4* __restrict__ a4, float4* __restrict__ b4, size_t n) {
    const f32x4v* a = (const f32x4v*)a4;
    f32x4v* b = (f32x4v*)b4;
    const size_t st = (size_t)gridDim.x * blockDim.x * COPY_UNROLL;
    for (size_t i = (size_t)blockIdx.x * blockDim.x * COPY_UNROLL + threadIdx.x; i < n; i += st) {
        f32x4v v[COPY_UNROLL];
#pragma unroll
        for (int u = 0; u < COPY_UNROLL; u++) {
            const size_t j = i + (size_t)u * blockDim.x;
            if (j < n) v[u] = __builtin_nontemporal_load(&a[j]);
        }
#pragma unroll
        for (int u = 0; u < COPY_UNROLL; u++) {
            const size_t j = i + (size_t)u * blockDim.x;
            if (j < n) __builtin_nontemporal_store(v[u], &b[j]);
        }
    }
}

void launch_copy(hipStream_t s, const float4* a, float4* b, size_t n) {
    if (!n) return;
    size_t blocks = std::min<size_t>((n + 256 * COPY_UNROLL - 1) / (256 * COPY_UNROLL), 32768);
    hipLaunchKernelGGL(k_copy, dim3((unsigned)blocks), dim3(256), 0, s, a, b, n);
    PBD_HIP(hipGetLastError());
}

__global__ void k_fill_u64(unsigned long long* p, size_t n, unsigned long long v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

void launch_fill_u64(hipStream_t s, unsigned long long* p, size_t n, unsigned long long v) {
    if (!n) return;
    hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
    PBD_HIP(hipGetLastError());
}

__global__ void k_min_sweeps(int* __restrict__ sweeps, const int* __restrict__ cap, int sw, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) sweeps[i] = min(cap[i], sw);
}

// Per-cloth sweep counts of a per-colour (global path) projection with budgets.
void launch_min_sweeps(hipStream_t s, int* sweeps, const int* cap, int sw, int n) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_min_sweeps, dim3((n + 255) / 256), dim3(256), 0, s, sweeps, cap, sw, n);
    PBD_HIP(hipGetLastError());
}

__global__ void k_fill_int(int* p, size_t n, int v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Fill kernel instead of hipMemsetAsync so that the substep can be captured
// into a graph.
void launch_fill_int(hipStream_t s, int* p, size_t n, int v) {
    if (!n) return;
    hipLaunchKernelGGL(k_fill_int, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
    PBD_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// LDS tile kernel.  One workgroup per tile: the tile's loaded vertices live
// in LDS as float4 for the whole launch; each colour phase processes the
// tile's prims of that colour that the phase's dependency cone needs (a
// prefix of the colour's list, the planner's per-sweep counts) and ends with
// a workgroup barrier.
//
// whole tiles (the tile is the entire cloth): optional fused advect on load,
//   sweeps until the residual check passes or the cap, fused integrate on store.
// partial tiles: exactly nsweeps sweeps on p_in -> p_out (interior only),
//   skipped (interior copied) when the cloth already met the tolerance in the
//   previous chunk; the residual of the owned prims is max-reduced into the
//   cloth's slot for this chunk.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double block_max(double v, double* red) {
    v = wave_max(v);
    int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    __syncthreads();
    if (lane == 0) red[wv] = v;
    __syncthreads();
    double m = 0.0;
    for (int i = 0; i < nw; i++) m = fmax(m, red[i]);
    return m;
}

// How many records or vertices a thread fetches ahead: distance records per
// colour phase, bending records per colour phase, and vertices when loading
// or storing a tile.  The distance-only kernel keeps them small so that it
// fits 64 VGPRs, which lets two 1024-thread workgroups share a CU (one
// loading or storing its tile while the other projects).
template <bool BEND>
struct TileFetch {
    static constexpr int dist = BEND ? 8 : 4;
    static constexpr int bend = 4;
    static constexpr int vert = 4;
};

// This thread's max relative residual over the tile's owned distance prims
// (records fetched PF at a time).
template <int PF>
__device__ __forceinline__ double own_residual(const float4* P, const uint32_t* __restrict__ own,
                                               const float* __restrict__ rest, int n) {
    const int T = blockDim.x;
    double m = 0.0;
    for (int j0 = threadIdx.x; j0 < n; j0 += PF * T) {
        uint32_t pr[PF];
        float rl[PF];
#pragma unroll
        for (int u = 0; u < PF; u++) {
            const int j = j0 + u * T;
            pr[u] = j < n ? own[j] : 0u;
            rl[u] = j < n ? rest[j] : 1.0f;
        }
#pragma unroll
        for (int u = 0; u < PF; u++)
            if (j0 + u * T < n) m = fmax(m, rel_residual(P[pr[u] & 0xffffu], P[pr[u] >> 16], rl[u]));
    }
    return m;
}

#define TILE_PARAMS                                                                                              \
    const int *__restrict__ order, const TileDev *__restrict__ tiles, const int *__restrict__ verts,              \
        const uint32_t *__restrict__ dist, const float *__restrict__ dist_rest, const uint32_t *__restrict__ bend, \
        const float *__restrict__ bend_rest, const uint32_t *__restrict__ own, const float *__restrict__ own_rest, \
        const int *__restrict__ cols, const ClothDev *__restrict__ cloths, const float *__restrict__ ktab_d,       \
        const float *__restrict__ ktab_b, float4 *__restrict__ x, float4 *__restrict__ v,                          \
        const float4 *__restrict__ p_in, float4 *__restrict__ p_out, unsigned long long *__restrict__ res_slots,   \
        int ncl, int *__restrict__ sweeps_out, const int *__restrict__ sweep_cap, SolveArgs a
#define TILE_ARGS                                                                                               \
    order, tiles, verts, dist, dist_rest, bend, bend_rest, own, own_rest, cols, cloths, ktab_d, ktab_b, x, v, p_in, \
        p_out, res_slots, ncl, sweeps_out, sweep_cap, a

// The hand-off of positions between chunks of one queued launch, whose
// workgroups may sit on different XCDs (each with its own L2): the tile that
// publishes a chunk writes its interior through to memory (store_coherent),
// every wave waits for its stores to complete, then one thread stores the
// flag; the next chunk loads its cone past the caches (load_coherent).  No
// cache-wide write-back or invalidation is needed, so the other work of the
// launch keeps its L2.  (An agent-scope acquire in one wave plus ordinary
// loads measured rare stale halo values.)
__device__ __forceinline__ float4 load_coherent(const float4* p) {
    const unsigned long long* w = reinterpret_cast<const unsigned long long*>(p);
    const unsigned long long lo = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    const unsigned long long hi = __hip_atomic_load(w + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    return make_float4(__uint_as_float((unsigned)lo), __uint_as_float((unsigned)(lo >> 32)),
                       __uint_as_float((unsigned)hi), __uint_as_float((unsigned)(hi >> 32)));
}
__device__ __forceinline__ void store_coherent(float4* p, float4 q) {
    unsigned long long* w = reinterpret_cast<unsigned long long*>(p);
    const unsigned long long lo = (unsigned long long)__float_as_uint(q.x) | ((unsigned long long)__float_as_uint(q.y) << 32);
    const unsigned long long hi = (unsigned long long)__float_as_uint(q.z) | ((unsigned long long)__float_as_uint(q.w) << 32);
    __hip_atomic_store(w, lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(w + 1, hi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Loads P[0, n) = p_in at the tile's loaded vertices vl (VERT_PREFETCH per
// thread per batch: indices first, then the gathers).
template <int VERT_PREFETCH>
__device__ __forceinline__ void tile_load_p(float4* P, const int* __restrict__ vl, const float4* __restrict__ p_in,
                                            int n) {
    const int T = blockDim.x;
    for (int i0 = threadIdx.x; i0 < n; i0 += VERT_PREFETCH * T) {
        int g[VERT_PREFETCH];
        float4 q[VERT_PREFETCH];
#pragma unroll
        for (int u = 0; u < VERT_PREFETCH; u++) g[u] = i0 + u * T < n ? vl[i0 + u * T] : -1;
#pragma unroll
        for (int u = 0; u < VERT_PREFETCH; u++) q[u] = p_in[max(g[u], 0)];
#pragma unroll
        for (int u = 0; u < VERT_PREFETCH; u++)
            if (g[u] >= 0) P[i0 + u * T] = q[u];
    }
}

// This thread's share of the owned distance prims' relative residuals on P:
// the max and the sum of squares in fixed point (the fallback rule's two
// measures, computed as the separate residual kernel computes them).
template <int PF>
__device__ __forceinline__ void own_resid_parts(const float4* P, const uint32_t* __restrict__ own,
                                                const float* __restrict__ rest, int n, double* m,
                                                unsigned long long* sq) {
    const int T = blockDim.x;
    for (int j0 = threadIdx.x; j0 < n; j0 += PF * T) {
        uint32_t pr[PF];
        float rl[PF];
#pragma unroll
        for (int u = 0; u < PF; u++) {
            const int j = j0 + u * T;
            pr[u] = j < n ? own[j] : 0u;
            rl[u] = j < n ? rest[j] : 1.0f;
        }
#pragma unroll
        for (int u = 0; u < PF; u++)
            if (j0 + u * T < n) {
                const float r = rel_res(P[pr[u] & 0xffffu], P[pr[u] >> 16], rl[u]);
                *m = fmax(*m, (double)r);
                *sq += sq_fixed(r);
            }
    }
}

constexpr int QUEUE_SPIN = 1 << 22;
constexpr int DECIDE_SPIN = 1 << 20;

// A bounded wait gave up: count it and raise the launch's abort flag, so
// that every other wait of the launch ends at once instead of timing out
// in turn (the results are then wrong, and the count says so).
__device__ __forceinline__ void queue_give_up(const TileQueue& q) {
    atomicAdd(q.timeouts, 1);
    __hip_atomic_store(&q.ctr[QC_ABORT], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ bool queue_aborted(const TileQueue& q) {
    return __hip_atomic_load(&q.ctr[QC_ABORT], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) != 0;
}

// Pause between two polls of a flag that bypass the caches, doubling from
// about 0.25 to 4 microseconds: with hundreds of workgroups waiting at once,
// polling flat out floods the memory lines that hold the flags and slows
// every other access of the launch.
__device__ __forceinline__ void poll_pause(int* backoff) {
    for (int z = 0; z < *backoff; z++) __builtin_amdgcn_s_sleep(8);
    *backoff = min(*backoff * 2, 16);
}

// The cloth's budget, keep flag, final p buffer and a zeroed sweep count,
// written through to memory by the tile that decides and complete before it
// publishes the decision (other tiles of the launch read the budget and add
// to the sweep count after seeing it).
__device__ __forceinline__ void record_decision(const TileQueue& q, const ClothDev& cd, int c, bool k,
                                                int* sweeps_out) {
    const int b = k ? q.warm_sweeps : q.full_sweeps;
    const int run = (min(b, q.iterations) + q.ts - 1) / q.ts;
    __hip_atomic_store(&q.keep[c], k ? 1 : 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(&q.cap[c], b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(&q.pfinal[c], cd.whole ? (q.start ^ 1) : (q.start ^ (run & 1)), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(&sweeps_out[c], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __builtin_amdgcn_s_waitcnt(0);
}

// The keep decision of the tile's cloth from this tile's residual parts
// (m0, sq0: unwarmed start; m1, sq1: with the prediction).  A single-tile
// cloth decides alone.  A multi-tile cloth's tiles add their sums and count
// themselves in; the last one applies the rule, records and publishes it
// and appends the cloth's later chunks (chunk-major, so that an item waits
// only on items before it); the others wait for the publication.  Every
// thread of the workgroup calls this; all get the same answer.
__device__ __forceinline__ bool cloth_decide(const TileQueue& q, const TileDev& td, const ClothDev& cd, int ncl,
                                             int* sweeps_out, double m0, double m1, unsigned long long sq0,
                                             unsigned long long sq1) {
    __shared__ unsigned long long acc[4];
    __shared__ int s_flag, s_base;
    const int c = td.cloth;
    if (threadIdx.x == 0) acc[0] = acc[1] = acc[2] = acc[3] = 0ull;
    for (int off = 32; off > 0; off >>= 1) {
        m0 = fmax(m0, __shfl_xor(m0, off));
        m1 = fmax(m1, __shfl_xor(m1, off));
        sq0 += __shfl_xor(sq0, off);
        sq1 += __shfl_xor(sq1, off);
    }
    __syncthreads();
    if ((threadIdx.x & 63) == 0) {
        // non-negative doubles order as their bit patterns
        atomicMax(&acc[0], (unsigned long long)__double_as_longlong(m0));
        atomicMax(&acc[1], (unsigned long long)__double_as_longlong(m1));
        atomicAdd(&acc[2], sq0);
        atomicAdd(&acc[3], sq1);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int flag;
        if (td.whole) {
            const bool k = q.rule == 1 ? acc[3] <= acc[2]
                                       : __longlong_as_double((long long)acc[1]) <= __longlong_as_double((long long)acc[0]);
            record_decision(q, cd, c, k, sweeps_out);
            flag = k ? 1 : 0;
        } else {
            if (q.rule == 1) {
                atomicAdd(&q.rs[2 * ncl + c], acc[2]);
                atomicAdd(&q.rs[3 * ncl + c], acc[3]);
            } else {
                atomicMax(&q.rs[c], acc[0]);
                atomicMax(&q.rs[ncl + c], acc[1]);
            }
            const int T = cd.tile_end - cd.tile_begin;
            __builtin_amdgcn_s_waitcnt(0);  // the sums are in before the tile counts itself in
            const int prev = __hip_atomic_fetch_add(&q.posted[c], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            if (prev == T - 1) {
                unsigned long long r[4];
                for (int k = 0; k < 4; k++)
                    r[k] = __hip_atomic_load(&q.rs[k * ncl + c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                const bool k = q.rule == 1 ? r[3] <= r[2]
                                           : __longlong_as_double((long long)r[1]) <= __longlong_as_double((long long)r[0]);
                record_decision(q, cd, c, k, sweeps_out);
                __hip_atomic_store(&q.decided[c * QC_DECIDED_STRIDE], k ? 2 : 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                flag = (k ? 1 : 0) | 2;
            } else {
                int d = 0, spins = 0, backoff = 1;
                while ((d = __hip_atomic_load(&q.decided[c * QC_DECIDED_STRIDE], __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_SYSTEM)) == 0) {
                    if (++spins > DECIDE_SPIN || ((spins & 63) == 0 && queue_aborted(q))) {
                        queue_give_up(q);
                        d = 1;
                        break;
                    }
                    poll_pause(&backoff);
                }
                flag = d - 1;
            }
        }
        s_flag = flag;
    }
    __syncthreads();
    const int flag = s_flag;
    if (flag & 2) {
        // this tile completed the decision: append the cloth's later chunks
        const int T = cd.tile_end - cd.tile_begin;
        const int b = (flag & 1) ? q.warm_sweeps : q.full_sweeps;
        const int more = (min(b, q.iterations) + q.ts - 1) / q.ts - 1;
        const int m = T * max(more, 0);
        if (m > 0) {
            if (threadIdx.x == 0) s_base = atomicAdd(&q.ctr[QC_DYN_TAIL], m);
            __syncthreads();
            const int base = s_base;
            for (int i = threadIdx.x; i < m; i += blockDim.x) {
                __hip_atomic_store(&q.dyn_tile[base + i], cd.tile_begin + i % T, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_SYSTEM);
                __hip_atomic_store(&q.dyn_chunk[base + i], 1 + i / T, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            }
            __builtin_amdgcn_s_waitcnt(0);  // written through and complete before the count is published
            __syncthreads();
            if (threadIdx.x == 0) {
                // publish in reservation order: the published count is a prefix
                int spins = 0;
                while (__hip_atomic_load(&q.ctr[QC_DYN_READY], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) != base) {
                    if (++spins > DECIDE_SPIN || ((spins & 63) == 0 && queue_aborted(q))) {
                        queue_give_up(q);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
                __hip_atomic_store(&q.ctr[QC_DYN_READY], base + m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            }
        }
        if (threadIdx.x == 0) __hip_atomic_fetch_add(&q.ctr[QC_DECIDED], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
    return (flag & 1) != 0;
}

// The tile kernel's body; BEND = false leaves the bending phases out (a plan
// with no bending prims).  DECIDE: a first chunk makes the warm start's keep
// decision itself (qd, see cloth_decide), and every budget is read from qd
// (the launch writes it; sweep_cap is unused).  Returns the cloth's sweep
// budget.  slot: the tile (< 0: none).
template <bool BEND, bool DECIDE = false>
__device__ __forceinline__ int tiles_body(TILE_PARAMS, int slot, float4* P, double* red,
                                          const TileQueue* qd = nullptr) {
    constexpr int PREFETCH = TileFetch<BEND>::dist;
    constexpr int BEND_PREFETCH = TileFetch<BEND>::bend;
    constexpr int VERT_PREFETCH = TileFetch<BEND>::vert;
    if (slot < 0) return 0;
    const TileDev td = tiles[slot];
    const ClothDev& cd = cloths[td.cloth];
    const int T = blockDim.x, tid = threadIdx.x;
    const int* vl = verts + td.load_off;

    // per-cloth sweep budget (warm start: short when the prediction was kept);
    // a cloth whose budget is spent is left where it is (its final buffer is
    // tracked per cloth), one that met the tolerance is carried forward
    const bool deciding = DECIDE && a.chunk == 0;
    int nsweeps = td.whole ? a.whole_nsweeps : a.nsweeps;
    int budget_c = 0x7fffffff;
    if (DECIDE && !deciding) {
        // written earlier in this launch by the cloth's decision
        budget_c = __hip_atomic_load(&qd->cap[td.cloth], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        nsweeps = min(nsweeps, budget_c - a.sweep_base);
    } else if (!DECIDE && sweep_cap) {
        budget_c = sweep_cap[td.cloth];
        nsweeps = min(nsweeps, budget_c - a.sweep_base);
    }
    const double tol = cd.tol >= 0.0 ? cd.tol : a.tol;
    bool carry = false;
    if (!td.whole && a.chunk > 0 && tol >= 0.0) {
        double prev = __longlong_as_double((long long)res_slots[(size_t)(a.chunk - 1) * ncl + td.cloth]);
        if (prev <= tol) carry = true;
    }
    if (!td.whole && nsweeps <= 0) {
        if (a.copy_spent) carry = true;
        else return budget_c;
    }
    // Gathers and scatters below go in batches of VERT_PREFETCH per thread:
    // indices first, then the dependent accesses, so each batch waits for
    // one memory round trip rather than one per vertex.
    if (carry) {
        for (int i0 = tid; i0 < td.n_int; i0 += VERT_PREFETCH * T) {
            int g[VERT_PREFETCH];
            float4 q[VERT_PREFETCH];
#pragma unroll
            for (int u = 0; u < VERT_PREFETCH; u++) g[u] = i0 + u * T < td.n_int ? vl[i0 + u * T] : -1;
#pragma unroll
            for (int u = 0; u < VERT_PREFETCH; u++) q[u] = p_in[max(g[u], 0)];
#pragma unroll
            for (int u = 0; u < VERT_PREFETCH; u++)
                if (g[u] >= 0) p_out[g[u]] = q[u];
        }
        return budget_c;
    }
    const int C = td.ncolors;
    const int* dc = cols + td.col_off;
    const int* bc = dc + C + 1;
    // partial tiles run, at sweep s of this launch, the prefix of each colour
    // list that the cone of the launch's last nsweeps sweeps needs (the index
    // is kept within the tile's plan_sweeps rows of counts), and load the
    // prefix of the vertex list that cone reads
    const int* dn = bc + C + 1;
    const int* bn = dn + td.plan_sweeps * C;
    const int* ln = bn + td.plan_sweeps * C;
    auto cone = [&](int n) { return (td.plan_sweeps && n < td.plan_sweeps) ? ln[n - 1] : td.n_load; };
    bool loaded = false;
    if (deciding) {
        // the unwarmed start of the kept budget's cone, its residuals, the
        // prediction added in LDS, the residuals again; then the decision
        const TileQueue& q = *qd;
        const int nk = min(nsweeps, q.warm_sweeps), nd = min(nsweeps, q.full_sweeps);
        const int lk = cone(nk);
        tile_load_p<VERT_PREFETCH>(P, vl, p_in, lk);
        __syncthreads();
        double m0 = 0.0, m1 = 0.0;
        unsigned long long sq0 = 0ull, sq1 = 0ull;
        own_resid_parts<PREFETCH>(P, own + td.own_off, own_rest + td.own_off, td.n_own, &m0, &sq0);
        __syncthreads();  // every read of the unwarmed P done before the prediction goes in
        for (int i0 = tid; i0 < lk; i0 += VERT_PREFETCH * T) {
            int g[VERT_PREFETCH];
            float4 dq[VERT_PREFETCH];
#pragma unroll
            for (int u = 0; u < VERT_PREFETCH; u++) g[u] = i0 + u * T < lk ? vl[i0 + u * T] : -1;
#pragma unroll
            for (int u = 0; u < VERT_PREFETCH; u++) dq[u] = a.delta[max(g[u], 0)];
#pragma unroll
            for (int u = 0; u < VERT_PREFETCH; u++)
                if (g[u] >= 0) {
                    const float4 q0 = P[i0 + u * T];
                    P[i0 + u * T] = make_float4(q0.x + dq[u].x, q0.y + dq[u].y, q0.z + dq[u].z, q0.w);
                }
        }
        __syncthreads();
        own_resid_parts<PREFETCH>(P, own + td.own_off, own_rest + td.own_off, td.n_own, &m1, &sq1);
        // (its barriers come after every read of P above)
        const bool k = cloth_decide(q, td, cd, ncl, sweeps_out, m0, m1, sq0, sq1);
        budget_c = k ? q.warm_sweeps : q.full_sweeps;
        nsweeps = k ? nk : nd;
        if (!k) {
            // discarded: the unwarmed start again, over the full budget's cone
            tile_load_p<VERT_PREFETCH>(P, vl, p_in, cone(nd));
        }
        loaded = true;
    }
    const int shift = max(0, td.plan_sweeps - nsweeps);
    const int n_load = cone(nsweeps);
    // the kept warm-start prediction enters here, on the substep's first load
    const bool add_delta = !loaded && a.delta != nullptr && a.keep[td.cloth] != 0;
    for (int i0 = tid; !loaded && i0 < n_load; i0 += VERT_PREFETCH * T) {
        int g[VERT_PREFETCH];
#pragma unroll
        for (int u = 0; u < VERT_PREFETCH; u++) g[u] = i0 + u * T < n_load ? vl[i0 + u * T] : -1;
        if (td.whole ? a.fused : a.advect_load) {
            float4 X[VERT_PREFETCH], V[VERT_PREFETCH];
#pragma unroll
            for (int u = 0; u < VERT_PREFETCH; u++) {
                X[u] = x[max(g[u], 0)];
                V[u] = v[max(g[u], 0)];
            }
#pragma unroll
            for (int u = 0; u < VERT_PREFETCH; u++)
                if (g[u] >= 0) {
                    float4 Vn;
                    P[i0 + u * T] = advect_one(cd, X[u], V[u], a.dt, &Vn);
                }
        } else if (add_delta) {
            float4 q[VERT_PREFETCH], dq[VERT_PREFETCH];
#pragma unroll
            for (int u = 0; u < VERT_PREFETCH; u++) {
                q[u] = p_in[max(g[u], 0)];
                dq[u] = a.delta[max(g[u], 0)];
            }
#pragma unroll
            for (int u = 0; u < VERT_PREFETCH; u++)
                if (g[u] >= 0) P[i0 + u * T] = make_float4(q[u].x + dq[u].x, q[u].y + dq[u].y, q[u].z + dq[u].z, q[u].w);
        } else if (a.coherent_load) {
            float4 q[VERT_PREFETCH];
#pragma unroll
            for (int u = 0; u < VERT_PREFETCH; u++) q[u] = load_coherent(p_in + max(g[u], 0));
#pragma unroll
            for (int u = 0; u < VERT_PREFETCH; u++)
                if (g[u] >= 0) P[i0 + u * T] = q[u];
        } else {
            float4 q[VERT_PREFETCH];
#pragma unroll
            for (int u = 0; u < VERT_PREFETCH; u++) q[u] = p_in[max(g[u], 0)];
#pragma unroll
            for (int u = 0; u < VERT_PREFETCH; u++)
                if (g[u] >= 0) P[i0 + u * T] = q[u];
        }
    }
    __syncthreads();

    const uint32_t* D = dist + td.dist_off;
    const float* DR = dist_rest + td.dist_off;
    const uint32_t* B = bend + 2 * (size_t)td.bend_off;
    const float* BR = bend_rest + td.bend_off;
    const bool coll = (cd.sphere_on || cd.box_on) && !a.no_coll;
    int sw = 0;
    double resid = -1.0;
    for (; sw < nsweeps;) {
        int gs = a.sweep_base + sw;
        int ks = gs < a.kmax ? gs : a.kmax - 1;
        float kd = ktab_d[cd.ktab + ks];
        float kb = ktab_b[cd.ktab + ks];
        const int* dnow = td.plan_sweeps ? dn + min(sw + shift, td.plan_sweeps - 1) * C : nullptr;
        const int* bnow = td.plan_sweeps ? bn + min(sw + shift, td.plan_sweeps - 1) * C : nullptr;
        for (int c = 0; c < C; c++) {
            // the prim records of a phase are fetched from global memory
            // in batches of PREFETCH per thread before any is processed, so
            // the phase waits for one round trip instead of one per prim
            const int je = dnow ? dc[c] + dnow[c] : dc[c + 1];
            for (int j0 = dc[c] + tid; j0 < je; j0 += PREFETCH * T) {
                uint32_t pr[PREFETCH];
                float rl[PREFETCH];
#pragma unroll
                for (int u = 0; u < PREFETCH; u++) {
                    const int j = j0 + u * T;
                    pr[u] = j < je ? D[j] : 0u;
                    rl[u] = j < je ? DR[j] : 0.0f;
                }
#pragma unroll
                for (int u = 0; u < PREFETCH; u++) {
                    if (j0 + u * T >= je) break;
                    const int i0 = pr[u] & 0xffffu, i1 = pr[u] >> 16;
                    float4 p0 = P[i0], p1 = P[i1];
                    project_distance(p0, p1, rl[u], kd);
                    P[i0] = p0;
                    P[i1] = p1;
                }
            }
            const int be = !BEND ? bc[c] : bnow ? bc[c] + bnow[c] : bc[c + 1];
            for (int j0 = bc[c] + tid; BEND && j0 < be; j0 += BEND_PREFETCH * T) {
                uint2 bw[BEND_PREFETCH];
                float br[BEND_PREFETCH];
#pragma unroll
                for (int u = 0; u < BEND_PREFETCH; u++) {
                    const int j = j0 + u * T;
                    bw[u] = j < be ? make_uint2(B[2 * j], B[2 * j + 1]) : make_uint2(0u, 0u);
                    br[u] = j < be ? BR[j] : 0.0f;
                }
#pragma unroll
                for (int u = 0; u < BEND_PREFETCH; u++) {
                    if (j0 + u * T >= be) break;
                    const int i0 = bw[u].x & 0xffffu, i1 = bw[u].x >> 16, i2 = bw[u].y & 0xffffu,
                              i3 = bw[u].y >> 16;
                    float4 q0 = P[i0], q1 = P[i1], q2 = P[i2], q3 = P[i3];
                    project_bending(q0, q1, q2, q3, br[u], kb);
                    P[i0] = q0;
                    P[i1] = q1;
                    P[i2] = q2;
                    P[i3] = q3;
                }
            }
            __syncthreads();
        }
        if (coll) {
            for (int i = tid; i < n_load; i += T) {
                float4 q = P[i];
                collide_point(q, cd.sphere_on, cd.sphere_c, cd.sphere_r, cd.box_on, cd.box_lo, cd.box_hi);
                P[i] = q;
            }
            __syncthreads();
        }
        sw++;
        if (td.whole && tol >= 0.0 && sw % a.check_every == 0) {
            resid = block_max(own_residual<PREFETCH>(P, own + td.own_off, own_rest + td.own_off, td.n_own), red);
            if (resid <= tol) break;
        }
    }
    // the owned prims' residual decides the tolerance stops; with no
    // tolerance (a fixed budget) nothing reads it, and it is not computed
    if (tol >= 0.0 && (!td.whole || resid < 0.0 || (sw % a.check_every) != 0))
        resid = block_max(own_residual<PREFETCH>(P, own + td.own_off, own_rest + td.own_off, td.n_own), red);
    if (tid == 0) {
        if (tol >= 0.0)
            atomicMax(&res_slots[(size_t)a.chunk * ncl + td.cloth], (unsigned long long)__double_as_longlong(resid));
        atomicMax(&sweeps_out[td.cloth], a.sweep_base + sw);
    }
    // integrate on store: a fused whole cloth, or the launch holding this
    // cloth's last sweep when the caller allows it (fixed budgets only)
    const int budget = min(budget_c, a.whole_nsweeps);
    const bool integ = (td.whole && a.fused) || (a.integrate_end && a.sweep_base + nsweeps >= budget);
    // a queued chunk with a further chunk of its cloth after it hands its stores on
    const bool handoff = a.queue_ts > 0 && !td.whole && budget_c > a.sweep_base + a.queue_ts;
    for (int i0 = tid; i0 < td.n_int; i0 += VERT_PREFETCH * T) {
        int g[VERT_PREFETCH];
#pragma unroll
        for (int u = 0; u < VERT_PREFETCH; u++) g[u] = i0 + u * T < td.n_int ? vl[i0 + u * T] : -1;
        if (integ) {
            float4 X[VERT_PREFETCH];
#pragma unroll
            for (int u = 0; u < VERT_PREFETCH; u++) X[u] = x[max(g[u], 0)];
#pragma unroll
            for (int u = 0; u < VERT_PREFETCH; u++)
                if (g[u] >= 0) {
                    const float4 Q = P[i0 + u * T];
                    v[g[u]] = make_float4((Q.x - X[u].x) / a.dt, (Q.y - X[u].y) / a.dt, (Q.z - X[u].z) / a.dt, 0.0f);
                    x[g[u]] = make_float4(Q.x, Q.y, Q.z, X[u].w);
                    p_out[g[u]] = Q;
                }
        } else if (handoff) {
#pragma unroll
            for (int u = 0; u < VERT_PREFETCH; u++)
                if (g[u] >= 0) store_coherent(p_out + g[u], P[i0 + u * T]);
        } else {
#pragma unroll
            for (int u = 0; u < VERT_PREFETCH; u++)
                if (g[u] >= 0) p_out[g[u]] = P[i0 + u * T];
        }
    }
    return budget_c;
}

// One tile per workgroup (n_items == nullptr: order[blockIdx.x]), or a
// fixed grid walking a list whose length the device wrote (*n_items): the
// workgroup takes entries blockIdx.x, blockIdx.x + gridDim.x, ... .  With
// gridDim.x a multiple of 8 every entry a workgroup takes sits at its own
// position mod 8, the XCD the list's layout meant it for.
template <bool BEND>
__device__ __forceinline__ void tiles_walk(TILE_PARAMS, const int* __restrict__ n_items, float4* P, double* red) {
    // one loop for both forms, so the body is inlined once
    const int n = n_items ? *n_items : (int)gridDim.x;
    for (int it = blockIdx.x; it < n; it += gridDim.x) {
        tiles_body<BEND>(TILE_ARGS, order[it], P, red);
        __syncthreads();  // the tile's stores have read LDS before the next tile loads
    }
}

// Plans without bending prims: at most 64 VGPRs (eight waves per SIMD), so
// two 1024-thread workgroups fit a CU when their tiles fit half its LDS.
__global__ __launch_bounds__(1024, 8) void k_tiles(TILE_PARAMS, const int* __restrict__ n_items) {
    extern __shared__ float4 P[];
    __shared__ double red[16];
    tiles_walk<false>(TILE_ARGS, n_items, P, red);
}

__global__ __launch_bounds__(1024) void k_tiles_bend(TILE_PARAMS, const int* __restrict__ n_items) {
    extern __shared__ float4 P[];
    __shared__ double red[16];
    tiles_walk<true>(TILE_ARGS, n_items, P, red);
}

// The tiles of a later chunk: the entries of `order` (the multi-tile launch
// order, laid out as order[j * 8 + g] for XCD group g, a group being a
// contiguous run of the plan's tiles; order_cloth holds each entry's cloth,
// -1 for padding) whose cloth's sweep budget cap exceeds min_cap.  They keep
// their order and are dealt again into 8 contiguous groups of equal size in
// the same layout, so neighbouring tiles still share an XCD and every XCD
// gets an equal share however the remaining cloths lie; the last group is
// padded with -1, and *n_out is the list's length.  One workgroup of 512
// threads: wave g ranks group g's entries with ballots, FOLLOW_BATCH entries
// per lane per round so that a round waits for two memory round trips.
constexpr int FOLLOW_BATCH = 8;

__global__ __launch_bounds__(512) void k_follow_list(const int* __restrict__ order, const int* __restrict__ order_cloth,
                                                     int n_order, const int* __restrict__ cap, int min_cap,
                                                     int* __restrict__ out, int* __restrict__ n_out) {
    __shared__ int counts[8];
    const int g = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int per_g = n_order / 8;
    // needed entries of this lane's batch starting at j0: the slots (or -1)
    auto batch = [&](int j0, int* slot) {
        int cl[FOLLOW_BATCH];
#pragma unroll
        for (int u = 0; u < FOLLOW_BATCH; u++) {
            const int j = j0 + u * 64 + lane;
            cl[u] = j < per_g ? order_cloth[j * 8 + g] : -1;
            slot[u] = j < per_g ? order[j * 8 + g] : -1;
        }
#pragma unroll
        for (int u = 0; u < FOLLOW_BATCH; u++)
            if (cl[u] < 0 || cap[cl[u]] <= min_cap) slot[u] = -1;
    };
    int count = 0;
    for (int j0 = 0; j0 < per_g; j0 += 64 * FOLLOW_BATCH) {
        int slot[FOLLOW_BATCH];
        batch(j0, slot);
#pragma unroll
        for (int u = 0; u < FOLLOW_BATCH; u++) count += __popcll(__ballot(slot[u] >= 0));
    }
    if (lane == 0) counts[g] = count;
    __syncthreads();
    int base = 0, total = 0;
    for (int k = 0; k < 8; k++) {
        base += k < g ? counts[k] : 0;
        total += counts[k];
    }
    const int per = (total + 7) / 8;  // entries per new group
    int r = base;
    for (int j0 = 0; j0 < per_g; j0 += 64 * FOLLOW_BATCH) {
        int slot[FOLLOW_BATCH];
        batch(j0, slot);
#pragma unroll
        for (int u = 0; u < FOLLOW_BATCH; u++) {
            const unsigned long long m = __ballot(slot[u] >= 0);
            if (slot[u] >= 0) {
                const int q = r + __popcll(m & ((1ull << lane) - 1ull));
                out[(q % per) * 8 + q / per] = slot[u];
            }
            r += __popcll(m);
        }
    }
    // pad: the positions of ranks total .. 8 * per - 1 (all in the last groups)
    for (int q = total + (int)threadIdx.x; q < 8 * per; q += (int)blockDim.x) out[(q % per) * 8 + q / per] = -1;
    if (threadIdx.x == 0) *n_out = 8 * per;
}

// Queue walk (see TileQueue).  A workgroup takes the next item; for a later
// chunk it first waits until every neighbour tile has finished the previous
// chunk (it read what they wrote, and writes what they read).  Deadlock-free
// whatever the number of resident workgroups: an item waits only for items
// taken before it, and a taken item is held by a running workgroup, so the
// oldest unfinished item can always proceed.  A tile whose cloth has a
// further chunk hands its chunk on as load_coherent describes (write-through
// stores, then the flag; the next chunk polls the flags and loads past the
// caches).  A wait gives up after QUEUE_SPIN polls and counts a timeout
// instead of hanging.

template <bool BEND>
__device__ __forceinline__ void tiles_queue(TILE_PARAMS, const TileQueue& q, float4* P, double* red) {
    __shared__ int s_item;
    const int n = *q.n_items;
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(q.head, 1);
        __syncthreads();
        const int idx = s_item;
        if (idx >= n) break;
        const int slot = q.item_tile[idx], c = q.item_chunk[idx];
        const TileDev& td = tiles[slot];
        // a later chunk of a cloth whose budget is already spent has nothing to
        // do (its neighbours publish no such chunk): no wait, no work
        if (c > 0 && sweep_cap && sweep_cap[td.cloth] <= c * q.ts) {
            __syncthreads();  // s_item read by all before the next take
            continue;
        }
        if (c > 0) {
            // poll relaxed; the chunk then loads its cone past the caches
            const int e = q.nbr_off[slot + 1];
            for (int k = q.nbr_off[slot] + (int)threadIdx.x; k < e; k += blockDim.x) {
                int spins = 0;
                while (__hip_atomic_load(&q.done[q.nbr[k]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) < c - 1) {
                    if (++spins > QUEUE_SPIN) {
                        atomicAdd(q.timeouts, 1);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(2);
                }
            }
            __syncthreads();
        }
        SolveArgs b = a;
        b.sweep_base = c * q.ts;
        b.nsweeps = min(q.ts, q.iterations - b.sweep_base);
        b.chunk = c;
        b.queue_ts = q.ts;
        if (c > 0) {
            b.delta = nullptr;
            b.keep = nullptr;
            b.advect_load = 0;
            b.coherent_load = 1;
        }
        const bool odd = ((q.start ^ c) & 1) != 0;
        tiles_body<BEND>(q.item_tile, tiles, verts, dist, dist_rest, bend, bend_rest, own, own_rest, cols, cloths,
                         ktab_d, ktab_b, x, v, odd ? q.pbuf1 : q.pbuf0, odd ? q.pbuf0 : q.pbuf1, res_slots, ncl,
                         sweeps_out, sweep_cap, b, slot, P, red);
        const bool more = !td.whole && sweep_cap && sweep_cap[td.cloth] > (c + 1) * q.ts;
        // every wave's write-through stores complete, then one thread publishes
        if (more) __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();  // also: s_item read by all before the next take
        if (more && threadIdx.x == 0) __hip_atomic_store(&q.done[slot], c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

__global__ __launch_bounds__(1024, 8) void k_tiles_queue(TILE_PARAMS, TileQueue q) {
    extern __shared__ float4 P[];
    __shared__ double red[16];
    tiles_queue<false>(TILE_ARGS, q, P, red);
}

__global__ __launch_bounds__(1024) void k_tiles_queue_bend(TILE_PARAMS, TileQueue q) {
    extern __shared__ float4 P[];
    __shared__ double red[16];
    tiles_queue<true>(TILE_ARGS, q, P, red);
}

// The queued launch that decides (see cloth_decide): a workgroup takes an
// appended later chunk when one is published, else the next fixed first
// chunk; with neither left it waits until every multi-tile cloth is decided
// and all appended items are taken.  Deadlock-free when every multi-tile
// cloth has fewer tiles than the resident workgroups: first chunks are
// taken in cloth order, so at most one cloth is partly taken, its waiting
// tiles leave a workgroup free for its remaining ones, and a later chunk
// waits only for items taken before it.
template <bool BEND>
__device__ __forceinline__ void tiles_decide_walk(TILE_PARAMS, const TileQueue& q, float4* P, double* red) {
    __shared__ int s_slot, s_chunk;
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(q.launches, 1);
    for (;;) {
        if (threadIdx.x < 64) {
            // wave 0 takes the item; lane 0 touches the counters
            const int lane = threadIdx.x;
            int slot = -1, ch = 0, spins = 0, backoff = 1;
            for (;;) {
                int h = 0, r = 0, st = 0;
                if (lane == 0) {
                    h = __hip_atomic_load(&q.ctr[QC_DYN_HEAD], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                    r = __hip_atomic_load(&q.ctr[QC_DYN_READY], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                    st = __hip_atomic_load(&q.ctr[QC_STAT], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                }
                h = __shfl(h, 0);
                r = __shfl(r, 0);
                const bool stat_left = __shfl(st, 0) < q.n_static;
                if (h < r) {
                    // the next appended item (written through before its count);
                    // while first chunks remain it is taken only if it can run
                    // at once, so that no workgroup sits on a chain meanwhile
                    const int hs = __hip_atomic_load(&q.dyn_tile[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                    const int hc = __hip_atomic_load(&q.dyn_chunk[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                    bool ok = true;
                    if (stat_left) {
                        const int e = q.nbr_off[hs + 1];
                        for (int k = q.nbr_off[hs] + lane; k < e; k += 64)
                            ok = ok && __hip_atomic_load(&q.finished[q.nbr[k]], __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_SYSTEM) >= hc;
                        ok = __all(ok);
                    }
                    if (ok) {
                        int got = 0;
                        if (lane == 0) got = atomicCAS(&q.ctr[QC_DYN_HEAD], h, h + 1) == h;
                        if (__shfl(got, 0)) {
                            slot = hs;
                            ch = hc;
                            break;
                        }
                        continue;
                    }
                }
                if (stat_left) {
                    int k = 0;
                    if (lane == 0) k = atomicAdd(&q.ctr[QC_STAT], 1);
                    k = __shfl(k, 0);
                    if (k < q.n_static) {
                        slot = q.stat_item[k];
                        break;
                    }
                    continue;
                }
                // nothing to take now; done once every multi-tile cloth has
                // decided (then every appended item is published) and taken
                int fin = 0;
                if (lane == 0 &&
                    __hip_atomic_load(&q.ctr[QC_DECIDED], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) == q.n_multi)
                    fin = __hip_atomic_load(&q.ctr[QC_DYN_HEAD], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) >=
                                  __hip_atomic_load(&q.ctr[QC_DYN_READY], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM)
                              ? 1
                              : 2;
                fin = __shfl(fin, 0);
                if (fin == 1) break;
                if (fin == 2) continue;
                int stop = 0;
                if (lane == 0 && (++spins > QUEUE_SPIN || ((spins & 63) == 0 && queue_aborted(q)))) {
                    queue_give_up(q);
                    stop = 1;
                }
                if (__shfl(stop, 0)) break;
                poll_pause(&backoff);
            }
            if (lane == 0) {
                s_slot = slot;
                s_chunk = ch;
            }
        }
        __syncthreads();
        const int slot = s_slot, c = s_chunk;
        if (slot < 0) break;
        if (c > 0) {
            // the neighbour tiles' previous chunk, polled by one thread in turn
            // (relaxed; the chunk then loads its cone past the caches)
            if (threadIdx.x == 0) {
                const int e = q.nbr_off[slot + 1];
                for (int k = q.nbr_off[slot]; k < e; k++) {
                    int spins = 0, backoff = 1;
                    while (__hip_atomic_load(&q.finished[q.nbr[k]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) < c) {
                        if (++spins > QUEUE_SPIN || ((spins & 63) == 0 && queue_aborted(q))) {
                            queue_give_up(q);
                            break;
                        }
                        poll_pause(&backoff);
                    }
                }
            }
            __syncthreads();
        }
        SolveArgs b = a;
        b.sweep_base = c * q.ts;
        b.nsweeps = min(q.ts, q.iterations - b.sweep_base);
        b.chunk = c;
        b.queue_ts = q.ts;
        if (c > 0) {
            b.delta = nullptr;
            b.keep = nullptr;
            b.advect_load = 0;
            b.coherent_load = 1;
        }
        const bool odd = ((q.start ^ c) & 1) != 0;
        const int budget = tiles_body<BEND, true>(order, tiles, verts, dist, dist_rest, bend, bend_rest, own,
                                                  own_rest, cols, cloths, ktab_d, ktab_b, x, v,
                                                  odd ? q.pbuf1 : q.pbuf0, odd ? q.pbuf0 : q.pbuf1, res_slots, ncl,
                                                  sweeps_out, nullptr, b, slot, P, red, &q);
        const bool more = !tiles[slot].whole && budget > (c + 1) * q.ts;
        if (more) __builtin_amdgcn_s_waitcnt(0);  // every wave's write-through stores complete
        __syncthreads();  // also: s_slot, s_chunk read by all before the next take
        if (more && threadIdx.x == 0)
            __hip_atomic_store(&q.finished[slot], c + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

__global__ __launch_bounds__(1024, 8) void k_tiles_decide(TILE_PARAMS, TileQueue q) {
    extern __shared__ float4 P[];
    __shared__ double red[16];
    tiles_decide_walk<false>(TILE_ARGS, q, P, red);
}

__global__ __launch_bounds__(1024) void k_tiles_decide_bend(TILE_PARAMS, TileQueue q) {
    extern __shared__ float4 P[];
    __shared__ double red[16];
    tiles_decide_walk<true>(TILE_ARGS, q, P, red);
}

// Block-wide exclusive prefix of a 0/1 flag over up to 1024 threads (16
// waves); returns this thread's offset and writes the total to *total.
__device__ __forceinline__ int block_flag_scan(bool f, int* wave_sum, int* total) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    const unsigned long long m = __ballot(f);
    if (lane == 0) wave_sum[wv] = (int)__popcll(m);
    __syncthreads();
    int off = 0, tot = 0;
    for (int w = 0; w < nw; w++) {
        off += w < wv ? wave_sum[w] : 0;
        tot += wave_sum[w];
    }
    __syncthreads();  // wave_sum read by all before its next use
    *total = tot;
    return off + (int)__popcll(m & ((1ull << lane) - 1ull));
}

__device__ __forceinline__ void queue_schedule_body(const int* __restrict__ order, int n_order,
                                                    const TileDev* __restrict__ tiles, int n_tiles,
                                                    const int* __restrict__ cap, int ts, int nchunks,
                                                    int* __restrict__ chain_tmp, int* __restrict__ bulk_tmp,
                                                    int* __restrict__ item_tile, int* __restrict__ item_chunk,
                                                    int* __restrict__ n_items, int* __restrict__ head,
                                                    int* __restrict__ done, int* wave_sum) {
    const int T = blockDim.x;
    for (int t = threadIdx.x; t < n_tiles; t += T) done[t] = -1;
    if (threadIdx.x == 0) *head = 0;
    int nc = 0, nb = 0;
    for (int j0 = 0; j0 < n_order; j0 += T) {
        const int j = j0 + (int)threadIdx.x;
        const int slot = j < n_order ? order[j] : -1;
        bool chain = false, bulk = false;
        if (slot >= 0) {
            const TileDev td = tiles[slot];
            chain = !td.whole && cap[td.cloth] > ts;
            bulk = !chain;
        }
        int tc, tb;
        const int oc = block_flag_scan(chain, wave_sum, &tc);
        const int ob = block_flag_scan(bulk, wave_sum, &tb);
        if (chain) chain_tmp[nc + oc] = slot;
        if (bulk) bulk_tmp[nb + ob] = slot;
        nc += tc;
        nb += tb;
    }
    __syncthreads();  // the lists written above, read below by other threads
    int base = 0;
    for (int k = 0; k < nchunks; k++) {
        for (int i = threadIdx.x; i < nc; i += T) {
            item_tile[base + i] = chain_tmp[i];
            item_chunk[base + i] = k;
        }
        base += nc;
        const int lo = (int)((long long)k * nb / nchunks), hi = (int)((long long)(k + 1) * nb / nchunks);
        for (int i = lo + (int)threadIdx.x; i < hi; i += T) {
            item_tile[base + i - lo] = bulk_tmp[i];
            item_chunk[base + i - lo] = 0;
        }
        base += hi - lo;
    }
    if (threadIdx.x == 0) *n_items = base;
}

__global__ __launch_bounds__(1024) void k_queue_schedule(const int* __restrict__ order, int n_order,
                                                         const TileDev* __restrict__ tiles, int n_tiles,
                                                         const int* __restrict__ cap, int ts, int nchunks,
                                                         int* __restrict__ chain_tmp, int* __restrict__ bulk_tmp,
                                                         int* __restrict__ item_tile, int* __restrict__ item_chunk,
                                                         int* __restrict__ n_items, int* __restrict__ head,
                                                         int* __restrict__ done) {
    __shared__ int wave_sum[16];
    queue_schedule_body(order, n_order, tiles, n_tiles, cap, ts, nchunks, chain_tmp, bulk_tmp, item_tile, item_chunk,
                        n_items, head, done, wave_sum);
}

// The keep decision and the substep's bookkeeping, then the schedule.
__global__ __launch_bounds__(1024) void k_queue_schedule_decide(
    QueueDecide d, const ClothDev* __restrict__ cl, int ncl, int* __restrict__ cap, int* __restrict__ sweeps,
    unsigned long long* __restrict__ res_slots, int n_slots, int* __restrict__ pfinal, int start, int iterations,
    const int* __restrict__ order, int n_order, const TileDev* __restrict__ tiles, int n_tiles, int ts, int nchunks,
    int* __restrict__ chain_tmp, int* __restrict__ bulk_tmp, int* __restrict__ item_tile, int* __restrict__ item_chunk,
    int* __restrict__ n_items, int* __restrict__ head, int* __restrict__ done) {
    __shared__ int wave_sum[16];
    for (int c = threadIdx.x; c < ncl; c += blockDim.x) {
        const bool k = ws_keep(d.rs, ncl, c, d.rule);
        d.keep[c] = k ? 1 : 0;
        const int b = k ? d.warm_sweeps : d.full_sweeps;
        cap[c] = b;
        sweeps[c] = 0;
        // the buffer the cloth's last chunk writes (a single-tile cloth: its one launch)
        const int run = (min(b, iterations) + ts - 1) / ts;
        pfinal[c] = cl[c].whole ? (start ^ 1) : (start ^ (run & 1));
    }
    for (int i = threadIdx.x; i < n_slots; i += blockDim.x) res_slots[i] = 0ull;
    __syncthreads();  // cap written above, read by every thread below
    queue_schedule_body(order, n_order, tiles, n_tiles, cap, ts, nchunks, chain_tmp, bulk_tmp, item_tile, item_chunk,
                        n_items, head, done, wave_sum);
}

#undef TILE_PARAMS
#undef TILE_ARGS

// Dynamic LDS the tile kernel may use on the current device: opts in to
// `bytes` and returns it, or returns the 64 KiB every launch may use without
// an opt-in when the runtime refuses it.  Every context on a device asks for
// the same amount (the device's per-workgroup maximum less the kernel's
// static LDS), so contexts never lower each other's limit.
size_t tiles_lds_optin(size_t bytes) {
    if (bytes <= 65536) return bytes;
    for (const void* k : {(const void*)k_tiles, (const void*)k_tiles_bend, (const void*)k_tiles_queue,
                          (const void*)k_tiles_queue_bend, (const void*)k_tiles_decide,
                          (const void*)k_tiles_decide_bend}) {
        const hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
        if (e != hipSuccess) {
            (void)hipGetLastError();  // the refusal is handled by the smaller limit
            return 65536;
        }
    }
    return bytes;
}

void launch_tiles(hipStream_t s, int nblocks, const int* order, const TileDev* tiles, const int* verts,
                  const uint32_t* dist, const float* dist_rest, const uint32_t* bend, const float* bend_rest,
                  const uint32_t* own, const float* own_rest, const int* cols, const ClothDev* cloths,
                  const float* ktab_d, const float* ktab_b, float4* x, float4* v, const float4* p_in,
                  float4* p_out, unsigned long long* res_slots, int ncl, int* sweeps_out, const int* sweep_cap,
                  const SolveArgs& a, int threads, size_t lds_bytes, bool with_bend, const int* n_items) {
    if (nblocks <= 0) return;
    hipLaunchKernelGGL(with_bend ? k_tiles_bend : k_tiles, dim3(nblocks), dim3(threads), lds_bytes, s, order, tiles,
                       verts, dist, dist_rest, bend, bend_rest, own, own_rest, cols, cloths, ktab_d, ktab_b, x, v,
                       p_in, p_out, res_slots, ncl, sweeps_out, sweep_cap, a, n_items);
    PBD_HIP(hipGetLastError());
}

void launch_tiles_queue(hipStream_t s, int nblocks, const TileDev* tiles, const int* verts, const uint32_t* dist,
                        const float* dist_rest, const uint32_t* bend, const float* bend_rest, const uint32_t* own,
                        const float* own_rest, const int* cols, const ClothDev* cloths, const float* ktab_d,
                        const float* ktab_b, float4* x, float4* v, unsigned long long* res_slots, int ncl,
                        int* sweeps_out, const int* sweep_cap, const SolveArgs& a, const TileQueue& q, int threads,
                        size_t lds_bytes, bool with_bend) {
    if (nblocks <= 0) return;
    // order, p_in and p_out of the common parameter list are unused: the
    // queue carries the items and the two buffers
    hipLaunchKernelGGL(with_bend ? k_tiles_queue_bend : k_tiles_queue, dim3(nblocks), dim3(threads), lds_bytes, s,
                       q.item_tile, tiles, verts, dist, dist_rest, bend, bend_rest, own, own_rest, cols, cloths,
                       ktab_d, ktab_b, x, v, (const float4*)q.pbuf0, q.pbuf1, res_slots, ncl, sweeps_out, sweep_cap,
                       a, q);
    PBD_HIP(hipGetLastError());
}

void launch_tiles_decide(hipStream_t s, int nblocks, const TileDev* tiles, const int* verts, const uint32_t* dist,
                         const float* dist_rest, const uint32_t* bend, const float* bend_rest, const uint32_t* own,
                         const float* own_rest, const int* cols, const ClothDev* cloths, const float* ktab_d,
                         const float* ktab_b, float4* x, float4* v, unsigned long long* res_slots, int ncl,
                         int* sweeps_out, const int* sweep_cap, const SolveArgs& a, const TileQueue& q, int threads,
                         size_t lds_bytes, bool with_bend) {
    if (nblocks <= 0) return;
    // order, p_in and p_out of the common parameter list are unused: the
    // queue carries the items and the two buffers
    hipLaunchKernelGGL(with_bend ? k_tiles_decide_bend : k_tiles_decide, dim3(nblocks), dim3(threads), lds_bytes, s,
                       q.stat_item, tiles, verts, dist, dist_rest, bend, bend_rest, own, own_rest, cols, cloths,
                       ktab_d, ktab_b, x, v, (const float4*)q.pbuf0, q.pbuf1, res_slots, ncl, sweeps_out, sweep_cap,
                       a, q);
    PBD_HIP(hipGetLastError());
}

void launch_queue_schedule(hipStream_t s, const int* order, int n_order, const TileDev* tiles, int n_tiles,
                           const int* cap, int ts, int nchunks, int* chain_tmp, int* bulk_tmp, int* item_tile,
                           int* item_chunk, int* n_items, int* head, int* done) {
    hipLaunchKernelGGL(k_queue_schedule, dim3(1), dim3(1024), 0, s, order, n_order, tiles, n_tiles, cap, ts, nchunks,
                       chain_tmp, bulk_tmp, item_tile, item_chunk, n_items, head, done);
    PBD_HIP(hipGetLastError());
}

void launch_queue_schedule_decide(hipStream_t s, const QueueDecide& d, const ClothDev* cl, int ncl, int* cap,
                                  int* sweeps, unsigned long long* res_slots, int n_slots, int* pfinal, int start,
                                  int iterations, const int* order, int n_order, const TileDev* tiles, int n_tiles,
                                  int ts, int nchunks, int* chain_tmp, int* bulk_tmp, int* item_tile, int* item_chunk,
                                  int* n_items, int* head, int* done) {
    hipLaunchKernelGGL(k_queue_schedule_decide, dim3(1), dim3(1024), 0, s, d, cl, ncl, cap, sweeps, res_slots, n_slots,
                       pfinal, start, iterations, order, n_order, tiles, n_tiles, ts, nchunks, chain_tmp, bulk_tmp,
                       item_tile, item_chunk, n_items, head, done);
    PBD_HIP(hipGetLastError());
}

void launch_follow_list(hipStream_t s, const int* order, const int* order_cloth, int n_order, const int* cap,
                        int min_cap, int* out, int* n_out) {
    hipLaunchKernelGGL(k_follow_list, dim3(1), dim3(512), 0, s, order, order_cloth, n_order, cap, min_cap, out,
                       n_out);
    PBD_HIP(hipGetLastError());
}

// Workgroups of the tile kernel one CU holds at once, for `threads` threads
// and `lds_bytes` of dynamic LDS.
int tiles_per_cu(bool with_bend, int threads, size_t lds_bytes) {
    int n = 0;
    PBD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, with_bend ? k_tiles_bend : k_tiles, threads,
                                                         lds_bytes));
    return std::max(1, n);
}

}  // namespace pbd
