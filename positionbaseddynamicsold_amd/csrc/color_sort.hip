// Constraint-graph colouring and a stable LDS radix-sort pass (gfx950).
//
// Colouring follows the reference's bulk-synchronous speculative scheme:
// AssignColorsGPU (PBD_cuda.cu:313-372) then ResolveConflictsGPU (:374-431),
// repeated until every prim is coloured.  Assign reads `color` and writes
// `prd` only for uncoloured prims; resolve reads `prd` and writes `color` of
// its own prim only, so each pass sees the previous pass's arrays and the
// result does not depend on scheduling.  Unlike the reference (a fixed 20000
// rounds, PBD_cuda.cu:508,826) the host stops as soon as no prim is left
// uncoloured.
#include "pbd_internal.hpp"

namespace pbd {

__global__ __launch_bounds__(256) void k_color_assign(const int2* __restrict__ adj_sn,
                                                      const int* __restrict__ adj,
                                                      const int* __restrict__ color,
                                                      int* __restrict__ prd, int np) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    if (color[p] >= 0) return;
    int2 sn = adj_sn[p];
    int c = -1, offset = 0;
    while (c < 0) {
        unsigned long long forbidden = 0ull;
        for (int i = sn.x; i < sn.x + sn.y; ++i) {
            int nc = color[adj[i]] - offset;
            if (nc >= 0 && nc < 64) forbidden |= (1ull << nc);
        }
        if (forbidden != ~0ull) {
            // first zero bit == log2((~x) & (x+1)) of the reference
            c = offset + __builtin_ctzll(~forbidden);
        } else {
            offset += 64;
        }
    }
    prd[p] = c;
}

// Keeps the proposal of every prim that has no conflicting neighbour of
// higher priority (more points, then lower id); the prims left uncoloured
// are counted into *uncolored with one atomic per workgroup (one per wave
// had the atomics on that one address take most of the round).
__global__ __launch_bounds__(256) void k_color_resolve(const int* __restrict__ npts,
                                                       const int2* __restrict__ adj_sn,
                                                       const int* __restrict__ adj,
                                                       int* __restrict__ color,
                                                       const int* __restrict__ prd, int np,
                                                       int* __restrict__ uncolored) {
    __shared__ int wave_left[4];
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    int left = 0;
    if (p < np && color[p] < 0) {
        int2 sn = adj_sn[p];
        int c = prd[p], conflict = 0, npt = npts[p];
        for (int i = sn.x; i < sn.x + sn.y; ++i) {
            int n = adj[i];
            if (prd[n] == c) {
                int nnpt = npts[n];
                if (nnpt > npt || (nnpt == npt && n < p)) {
                    conflict = 1;
                    break;
                }
            }
        }
        if (conflict) left = 1;
        else color[p] = c;
    }
    const unsigned long long m = __ballot(left);
    if ((threadIdx.x & 63) == 0) wave_left[threadIdx.x >> 6] = (int)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        const int tot = wave_left[0] + wave_left[1] + wave_left[2] + wave_left[3];
        if (tot) atomicAdd(uncolored, tot);
    }
}

int launch_color_round(hipStream_t s, const int* npts, const int* adj_sn, const int* adj, int* color, int* prd,
                       int* uncolored, int np) {
    int nb = (np + 255) / 256;
    if (nb == 0) return 0;
    hipLaunchKernelGGL(k_color_assign, dim3(nb), dim3(256), 0, s, (const int2*)adj_sn, adj, color, prd, np);
    hipLaunchKernelGGL(k_color_resolve, dim3(nb), dim3(256), 0, s, npts, (const int2*)adj_sn, adj, color, prd,
                       np, uncolored);
    PBD_HIP(hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------------------
// Stable LSD radix pass over (key, value) uint32 pairs, 8-bit digits.
// Blocks own 4096 consecutive elements.  hist[d * nb + b] counts digit d in
// block b; an exclusive scan over that digit-major array gives every
// (digit, block) its first output slot, and the scatter ranks equal digits
// inside a block in input order (wave match + per-wave counts), so the pass
// is stable.
// ---------------------------------------------------------------------------
constexpr int RB = 4096;

__global__ __launch_bounds__(256) void k_radix_hist(const uint32_t* __restrict__ keys, int n,
                                                    int shift, int* __restrict__ hist, int nb) {
    __shared__ int h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    int base = blockIdx.x * RB;
    for (int i = threadIdx.x; i < RB; i += 256) {
        int g = base + i;
        if (g < n) atomicAdd(&h[(keys[g] >> shift) & 255u], 1);
    }
    __syncthreads();
    hist[threadIdx.x * nb + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(256) void k_radix_scatter(const uint32_t* __restrict__ kin,
                                                       const uint32_t* __restrict__ vin,
                                                       uint32_t* __restrict__ kout,
                                                       uint32_t* __restrict__ vout, int n, int shift,
                                                       const int* __restrict__ offs, int nb) {
    __shared__ int base[256];
    __shared__ int wcnt[4][256];
    int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    base[t] = offs[t * nb + blockIdx.x];
    __syncthreads();
    for (int r = 0; r < RB / 256; r++) {
        int i = blockIdx.x * RB + r * 256 + t;
        bool valid = i < n;
        uint32_t k = valid ? kin[i] : 0u, v = valid ? vin[i] : 0u;
        int d = valid ? (int)((k >> shift) & 255u) : -1;
        for (int j = lane; j < 256; j += 64) wcnt[wv][j] = 0;
        // lanes with this lane's digit: one ballot per digit bit, so the cost
        // does not depend on how many distinct digits the wave holds
        unsigned long long m = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const bool bit = (d >> b) & 1;
            const unsigned long long bb = __ballot(valid && bit);
            m &= bit ? bb : ~bb;
        }
        int rank = 0;
        if (valid) {
            rank = __popcll(m & __lanemask_lt());
            if (rank == 0) wcnt[wv][d] = __popcll(m);  // the digit's first lane
        }
        __syncthreads();
        if (valid) {
            int pos = base[d] + rank;
            for (int w2 = 0; w2 < wv; w2++) pos += wcnt[w2][d];
            kout[pos] = k;
            vout[pos] = v;
        }
        __syncthreads();
        base[t] += wcnt[0][t] + wcnt[1][t] + wcnt[2][t] + wcnt[3][t];
        __syncthreads();
    }
}

// Stable sort of (key, value) pairs on bits [0, bits) of the key.  Uses
// (k2, v2) as the ping-pong buffer; the result ends in (k, v).  `hist` holds
// radix_hist_ints(n) ints: digit counts | their exclusive scan (multi-block)
// | the scan's block sums.
void radix_sort_pairs(hipStream_t s, uint32_t* k, uint32_t* v, uint32_t* k2, uint32_t* v2, int n,
                      int bits, int* hist) {
    if (n <= 0) return;
    int nb = (n + RB - 1) / RB;
    int passes = (bits + 7) / 8;
    int* counts = hist;
    int* offs = hist + (size_t)256 * nb;
    int* sums = hist + (size_t)512 * nb;
    uint32_t *ki = k, *vi = v, *ko = k2, *vo = v2;
    for (int ps = 0; ps < passes; ps++) {
        hipLaunchKernelGGL(k_radix_hist, dim3(nb), dim3(256), 0, s, ki, n, ps * 8, counts, nb);
        device_scan(s, counts, offs, 256 * nb, sums);
        hipLaunchKernelGGL(k_radix_scatter, dim3(nb), dim3(256), 0, s, ki, vi, ko, vo, n, ps * 8, offs, nb);
        std::swap(ki, ko);
        std::swap(vi, vo);
    }
    if (ki != k) {
        PBD_HIP(hipMemcpyAsync(k, ki, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, s));
        PBD_HIP(hipMemcpyAsync(v, vi, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, s));
    }
    PBD_HIP(hipGetLastError());
}

}  // namespace pbd
