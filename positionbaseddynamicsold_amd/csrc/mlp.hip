// Learned warm start: a per-vertex MLP 22 -> 64 -> 64 -> 3 (GELU, tanh form)
// with BF16 weights and activations on MFMA (v_mfma_f32_16x16x16_bf16) and
// FP32 accumulation; forward, backward and Adam kernels.  No counterpart in
// the reference.
//
// Feature vector of vertex i (cloth mean rest length h, predicted p, previous x):
//   [0:3]   (p_i - x_i) / h
//   [3:21]  for its first 6 distance prims (prim order), neighbour j:
//           ((|e| - L) / L) * e / |e|,  e = p_j - p_i   (zeros when absent)
//   [21]    1 if the vertex is pinned (w = 0), else 0
// Output y (3): the predicted Gauss-Seidel correction divided by h.
//
// Parameter layout (FP32 master copy, row-major W[out][in]):
//   W1[64][22] b1[64] W2[64][64] b2[64] W3[3][64] b3[3]
#include <algorithm>
#include <cstddef>
#include <map>
#include <mutex>
#include <cmath>

#include "pbd_internal.hpp"

namespace pbd {

// Makes LDS writes of this wave visible to its other lanes.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// bf16 round-to-nearest-even (v_cvt_pk_bf16_f32; NaN stays NaN)
__device__ __forceinline__ uint16_t f2bf(float x) { return __builtin_bit_cast(uint16_t, (__bf16)x); }

typedef float f32v2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16v2 __attribute__((ext_vector_type(2)));
// Round-to-nearest-even BF16 of four floats; the two-element conversions map
// to one v_cvt_pk_bf16_f32 per pair.
__device__ __forceinline__ bf16x4 pack_bf16(float a, float b, float c, float d) {
    const bf16v2 lo = __builtin_convertvector((f32v2){a, b}, bf16v2);
    const bf16v2 hi = __builtin_convertvector((f32v2){c, d}, bf16v2);
    return __builtin_bit_cast(bf16x4, make_uint2(__builtin_bit_cast(uint32_t, lo), __builtin_bit_cast(uint32_t, hi)));
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// GELU in its tanh form, 0.5 x (1 + tanh(u)) with u = sqrt(2/pi) (x + 0.044715 x^3),
// evaluated as the identical x * sigmoid(2u) on the hardware exp2 and
// reciprocal, with the constants folded so that exp(-2u) = 2^(x (B + A x^2)):
// five VALU operations and two transcendentals per value (the inference
// kernel's bound).  Saturates to 0 / x without clamping.  Training and
// inference share it.  dy/dx = s + x s (1 - s) 2u'(x).
__device__ __forceinline__ void gelu(float x, float* y, float* dy) {
    const float c2 = 1.5957691216f;  // 2 sqrt(2/pi)
    const float k = 0.044715f;
    const float B = -2.3022082f;     // -2 sqrt(2/pi) log2(e)
    const float A = -0.10294324f;    // 0.044715 B
    const float x2 = x * x;
    const float e = __builtin_amdgcn_exp2f(x * __builtin_fmaf(A, x2, B));  // exp(-2u)
    const float sg = __builtin_amdgcn_rcpf(1.0f + e);
    *y = x * sg;
    if (dy) *dy = sg + x * sg * (1.0f - sg) * c2 * (1.0f + 3.0f * k * x2);
}

// Standardised feature as the network sees it (training and inference);
// rsd = 1 / sd rounded to FP32.
__device__ __forceinline__ float norm_feature(float f, float mu, float rsd) {
    return fminf(fmaxf((f - mu) * rsd, -8.0f), 8.0f);
}

// One wave: C[16 x 16*nt] = A[16 x K] * B, with A row-major (lda) and B given
// as Bt = B^T row-major [16*nt x K] (ldb), all BF16 in LDS.  acc[t][i] is
// C[4*(lane/16) + i][16*t + lane%16].
__device__ __forceinline__ void mm16(const uint16_t* A, int lda, const uint16_t* Bt, int ldb, int K, int nt,
                                     f32x4* acc) {
    const int l = threadIdx.x & 63, r = l & 15, q = l >> 4;
    for (int t = 0; t < nt; t++) acc[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    for (int k0 = 0; k0 < K; k0 += 16) {
        bf16x4 a = *(const bf16x4*)(A + r * lda + k0 + 4 * q);
        for (int t = 0; t < nt; t++) {
            bf16x4 b = *(const bf16x4*)(Bt + (t * 16 + r) * ldb + k0 + 4 * q);
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[t], 0, 0, 0);
        }
    }
}

// Weights as the kernels read them (BF16, K padded to multiples of 16).
struct alignas(16) MlpW {
    uint16_t w1[64 * 32];    // W1[out][in], in padded 22 -> 32
    uint16_t w2[64 * 64];    // W2[out][in]
    uint16_t w3[16 * 64];    // W3[out][in], out padded 3 -> 16
    uint16_t w2t[64 * 64];   // W2^T [in][out]
    // inference copies: row o of layer n holds W_n[o][16s + 4q + j] at
    // 8q + 4s + j (layer 1) / 16q + 4s + j (layers 2, 3), so that lane quad q
    // reads all its K-steps of a row with one 16- or 32-byte load
    uint16_t i1[64 * 32];
    uint16_t i2[64 * 64];
    uint16_t i3[16 * 64];
    float b1[64], b2[64], b3[16];
};

constexpr int P_W1 = 0, P_B1 = P_W1 + 64 * 22, P_W2 = P_B1 + 64, P_B2 = P_W2 + 64 * 64, P_W3 = P_B2 + 64,
              P_B3 = P_W3 + 3 * 64, NPARAM = P_B3 + 3;
static_assert(NPARAM == 5827, "parameter count");

__global__ __launch_bounds__(256) void k_mlp_pack(const float* __restrict__ prm, MlpW* __restrict__ W) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 64 * 32) {
        int o = i / 32, k = i % 32;
        W->w1[i] = k < 22 ? f2bf(prm[P_W1 + o * 22 + k]) : 0;
    }
    if (i < 64 * 64) {
        int o = i / 64, k = i % 64;
        uint16_t v = f2bf(prm[P_W2 + o * 64 + k]);
        W->w2[i] = v;
        W->w2t[k * 64 + o] = v;
    }
    if (i < 16 * 64) {
        int o = i / 64, k = i % 64;
        W->w3[i] = o < 3 ? f2bf(prm[P_W3 + o * 64 + k]) : 0;
    }
    if (i < 64 * 32) {
        int o = i / 32, c = i % 32, q = c / 8, st = (c % 8) / 4, j = c % 4, k = 16 * st + 4 * q + j;
        W->i1[i] = k < 22 ? f2bf(prm[P_W1 + o * 22 + k]) : 0;
    }
    if (i < 64 * 64) {
        int o = i / 64, c = i % 64, q = c / 16, st = (c % 16) / 4, j = c % 4, k = 16 * st + 4 * q + j;
        W->i2[i] = f2bf(prm[P_W2 + o * 64 + k]);
        if (o < 16) W->i3[i] = o < 3 ? f2bf(prm[P_W3 + o * 64 + k]) : 0;
    }
    if (i < 64) {
        W->b1[i] = prm[P_B1 + i];
        W->b2[i] = prm[P_B2 + i];
    }
    if (i < 16) W->b3[i] = i < 3 ? prm[P_B3 + i] : 0.0f;
}

// Per-wave scratch of the forward pass.
struct alignas(16) FwdLds {
    uint16_t xs[16 * 32];
    uint16_t a1[16 * 64];
    uint16_t a2[16 * 64];
    float h1[16 * 64];
    float h2[16 * 64];
    float y[16 * 16];
};

__device__ __forceinline__ void load_weights(MlpW* sW, const MlpW* __restrict__ gW) {
    const uint4* s = (const uint4*)gW;
    uint4* d = (uint4*)sW;
    for (int i = threadIdx.x; i < (int)(sizeof(MlpW) / 16); i += blockDim.x) d[i] = s[i];
    __syncthreads();
}

// Forward of 16 rows whose BF16 inputs are already in F.xs.  Fills h1/a1/h2/a2
// and y (16 x 16, columns 0..2 meaningful).
__device__ void forward16(const MlpW& W, FwdLds& F) {
    const int l = threadIdx.x & 63, r = l & 15, q = l >> 4;
    f32x4 acc[4];
    mm16(F.xs, 32, W.w1, 32, 32, 4, acc);
    for (int t = 0; t < 4; t++)
        for (int i = 0; i < 4; i++) {
            int row = 4 * q + i, col = 16 * t + r;
            float h = acc[t][i] + W.b1[col], a;
            gelu(h, &a, nullptr);
            F.h1[row * 64 + col] = h;
            F.a1[row * 64 + col] = f2bf(a);
        }
    wave_sync();
    mm16(F.a1, 64, W.w2, 64, 64, 4, acc);
    for (int t = 0; t < 4; t++)
        for (int i = 0; i < 4; i++) {
            int row = 4 * q + i, col = 16 * t + r;
            float h = acc[t][i] + W.b2[col], a;
            gelu(h, &a, nullptr);
            F.h2[row * 64 + col] = h;
            F.a2[row * 64 + col] = f2bf(a);
        }
    wave_sync();
    mm16(F.a2, 64, W.w3, 64, 64, 1, acc);
    for (int i = 0; i < 4; i++) {
        int row = 4 * q + i, col = r;
        F.y[row * 16 + col] = acc[0][i] + W.b3[col];
    }
    wave_sync();
}

// ---------------------------------------------------------------------------
// batch forward on stored samples (evaluation / tests)
// ---------------------------------------------------------------------------
constexpr int FWD_WAVES = 4;

__global__ __launch_bounds__(256) void k_mlp_forward(const MlpW* __restrict__ gW, const float* __restrict__ X,
                                                     int n, float* __restrict__ Y) {
    __shared__ MlpW W;
    __shared__ FwdLds F[FWD_WAVES];
    load_weights(&W, gW);
    const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
    FwdLds& f = F[wv];
    int base = (blockIdx.x * FWD_WAVES + wv) * 16;
    if (base >= n) return;
    for (int e = l; e < 16 * 32; e += 64) {
        int row = e / 32, k = e % 32, s = base + row;
        f.xs[e] = (k < 22 && s < n) ? f2bf(X[(size_t)s * 22 + k]) : 0;
    }
    wave_sync();
    forward16(W, f);
    if (l < 48) {
        int row = l / 3, o = l % 3, s = base + row;
        if (s < n) Y[(size_t)s * 3 + o] = f.y[row * 16 + o];
    }
}

// ---------------------------------------------------------------------------
// training: forward + backward over 16 samples per wave; per-wave parameter
// gradients written to partial[wave][NPARAM+1] (last slot: loss)
// ---------------------------------------------------------------------------
struct alignas(16) BwdLds {
    uint16_t xT[32 * 16];     // X^T [in][sample]
    uint16_t a1T[64 * 16];    // A1^T
    uint16_t a2T[64 * 16];    // A2^T
    uint16_t dyT[16 * 16];    // dY^T [out16][sample]
    uint16_t dh2[16 * 64];    // dH2 [sample][o]
    uint16_t dh2T[64 * 16];   // dH2^T
    uint16_t dh1T[64 * 16];   // dH1^T
    float dh2f[16 * 64];
    float dh1f[16 * 64];
    float dy[16 * 4];
};

constexpr int TRAIN_WAVES = 2;

__global__ __launch_bounds__(128) void k_mlp_train(const MlpW* __restrict__ gW, const float* __restrict__ X,
                                                   const float* __restrict__ T, const int* __restrict__ idx,
                                                   int nbatch, float inv_total, float* __restrict__ partial) {
    __shared__ MlpW W;
    __shared__ FwdLds F[TRAIN_WAVES];
    __shared__ BwdLds B[TRAIN_WAVES];
    load_weights(&W, gW);
    const int wv = threadIdx.x >> 6, l = threadIdx.x & 63, r = l & 15, q = l >> 4;
    FwdLds& f = F[wv];
    BwdLds& b = B[wv];
    const int gw = blockIdx.x * TRAIN_WAVES + wv;
    const int base = gw * 16;
    float* out = partial + (size_t)gw * (NPARAM + 1);
    if (base >= nbatch) return;
    // gather the 16 samples
    for (int e = l; e < 16 * 32; e += 64) {
        int row = e / 32, k = e % 32, s = base + row;
        float v = (k < 22 && s < nbatch) ? X[(size_t)idx[s] * 22 + k] : 0.0f;
        uint16_t hv = f2bf(v);
        f.xs[e] = hv;
        b.xT[k * 16 + row] = hv;
    }
    wave_sync();
    forward16(W, f);
    // dY = (y - t) / N (loss = sum 0.5 |y - t|^2 / N)
    float lossp = 0.0f;
    if (l < 48) {
        int row = l / 3, o = l % 3, s = base + row;
        float dy = 0.0f;
        if (s < nbatch) {
            float e = f.y[row * 16 + o] - T[(size_t)idx[s] * 3 + o];
            lossp = 0.5f * e * e * inv_total;
            dy = e * inv_total;
        }
        b.dy[row * 4 + o] = dy;
    }
    wave_sync();
    for (int e = l; e < 16 * 16; e += 64) {
        int o = e / 16, s = e % 16;
        b.dyT[e] = o < 3 ? f2bf(b.dy[s * 4 + o]) : 0;
    }
    // transposed activations for the weight-gradient products
    for (int e = l; e < 16 * 64; e += 64) {
        int row = e / 64, c = e % 64;
        b.a1T[c * 16 + row] = f.a1[e];
        b.a2T[c * 16 + row] = f.a2[e];
    }
    wave_sync();
    // dA2 = dY W3 (VALU, 3 terms), dH2 = dA2 * gelu'(H2)
    for (int e = l; e < 16 * 64; e += 64) {
        int row = e / 64, c = e % 64;
        int s = base + row;
        float da = 0.0f;
        for (int o = 0; o < 3; o++) da += b.dy[row * 4 + o] * bf2f(W.w3[o * 64 + c]);
        float g, gd;
        gelu(f.h2[e], &g, &gd);
        float d = s < nbatch ? da * gd : 0.0f;
        b.dh2f[e] = d;
        uint16_t hv = f2bf(d);
        b.dh2[e] = hv;
        b.dh2T[c * 16 + row] = hv;
    }
    wave_sync();
    // dA1 = dH2 W2 (MFMA with B^T = W2^T stored as w2t), dH1 = dA1 * gelu'(H1)
    {
        f32x4 acc[4];
        mm16(b.dh2, 64, W.w2t, 64, 64, 4, acc);
        for (int t = 0; t < 4; t++)
            for (int i = 0; i < 4; i++) {
                int row = 4 * q + i, col = 16 * t + r;
                float g, gd;
                gelu(f.h1[row * 64 + col], &g, &gd);
                float d = (base + row < nbatch) ? acc[t][i] * gd : 0.0f;
                b.dh1f[row * 64 + col] = d;
                b.dh1T[col * 16 + row] = f2bf(d);
            }
    }
    wave_sync();
    // weight gradients: dW = dH^T A (MFMA, K = the 16 samples)
    {
        f32x4 acc[4];
        // dW3 [3 x 64]
        mm16(b.dyT, 16, b.a2T, 16, 16, 4, acc);
        for (int t = 0; t < 4; t++)
            for (int i = 0; i < 4; i++) {
                int o = 4 * q + i, c = 16 * t + r;
                if (o < 3) out[P_W3 + o * 64 + c] = acc[t][i];
            }
        // dW2 [64 x 64]: four row tiles of o
        for (int ot = 0; ot < 4; ot++) {
            mm16(b.dh2T + ot * 16 * 16, 16, b.a1T, 16, 16, 4, acc);
            for (int t = 0; t < 4; t++)
                for (int i = 0; i < 4; i++) {
                    int o = ot * 16 + 4 * q + i, c = 16 * t + r;
                    out[P_W2 + o * 64 + c] = acc[t][i];
                }
        }
        // dW1 [64 x 22]: K padded 32
        for (int ot = 0; ot < 4; ot++) {
            mm16(b.dh1T + ot * 16 * 16, 16, b.xT, 16, 16, 2, acc);
            for (int t = 0; t < 2; t++)
                for (int i = 0; i < 4; i++) {
                    int o = ot * 16 + 4 * q + i, c = 16 * t + r;
                    if (c < 22) out[P_W1 + o * 22 + c] = acc[t][i];
                }
        }
    }
    // bias gradients (FP32 sums over the 16 samples) and the loss
    {
        float s1 = 0.0f, s2 = 0.0f;
        for (int row = 0; row < 16; row++) {
            s1 += b.dh1f[row * 64 + l];
            s2 += b.dh2f[row * 64 + l];
        }
        out[P_B1 + l] = s1;
        out[P_B2 + l] = s2;
        if (l < 3) {
            float s3 = 0.0f;
            for (int row = 0; row < 16; row++) s3 += b.dy[row * 4 + l];
            out[P_B3 + l] = s3;
        }
        for (int off = 32; off > 0; off >>= 1) lossp += __shfl_xor(lossp, off);
        if (l == 0) out[NPARAM] = lossp;
    }
}

// Sum of the per-wave partials: a block takes 16 parameters x 16 wave
// slices; slice j adds waves j, j + 16, ... in order and the 16 slice sums
// are added in slice order (deterministic).
__global__ __launch_bounds__(256) void k_mlp_reduce(const float* __restrict__ partial, int nw,
                                                    float* __restrict__ grad) {
    __shared__ float red[16][17];
    const int pi = threadIdx.x & 15, sl = threadIdx.x >> 4;
    const int i = blockIdx.x * 16 + pi;
    float acc = 0.0f;
    if (i <= NPARAM)
        for (int w = sl; w < nw; w += 16) acc += partial[(size_t)w * (NPARAM + 1) + i];
    red[sl][pi] = acc;
    __syncthreads();
    if (sl == 0 && i <= NPARAM) {
        float t = 0.0f;
        for (int k = 0; k < 16; k++) t += red[k][pi];
        grad[i] = t;
    }
}

// Adam (Kingma & Ba 2015) on the FP32 master parameters; c1 = 1 - b1^t, c2 = 1 - b2^t.
__global__ __launch_bounds__(256) void k_adam(float* __restrict__ w, const float* __restrict__ g,
                                              float* __restrict__ m, float* __restrict__ v, int n, float lr,
                                              float b1, float b2, float eps, float c1, float c2) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float gi = g[i];
    float mi = b1 * m[i] + (1.0f - b1) * gi;
    float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    float mh = mi / c1, vh = vi / c2;
    w[i] = w[i] - lr * mh / (sqrtf(vh) + eps);
}

// ---------------------------------------------------------------------------
// warm start inside the substep
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// Inference in two kernels.
//
// k_ws_features: one thread per vertex gathers its 22 features, standardises
// them and writes them as one 64-byte BF16 row laid out in MFMA operand order:
// 16-byte chunk q holds features 4q..4q+3 then 16+4q..16+4q+3 (zero past 21),
// so lane (r, q) of the network kernel loads both K-steps of its A operand
// for vertex r with a single 16-byte load.  Also writes the output scale
// h (0 for pinned vertices, and with a gate > 0 for vertices whose one-ring
// prims are all within gate of their rest length: no correction is predicted
// where no constraint is violated) and, when asked, the raw FP32 features.
//
// k_ws_mlp: persistent; the weights stay in LDS, each wave takes 16 vertices
// at a time with the next group's features already in flight, and writes
// delta = h * (y * sdy) straight from the accumulators.
// ---------------------------------------------------------------------------
// Features of vertex v: the standardised BF16 row in MFMA operand order
// (b[4q..4q+3] = 16-byte chunk q), the output scale (returned) and, when
// asked, the raw FP32 features.  s_mu / s_rsd: feature mean and 1 / sd.
// ADVECT: p is not given; the predicted positions of v and its neighbours are
// computed here from x and vel (the same advect_one as the advect kernel) and
// v's is written to p_out.
template <bool ADVECT>
__device__ __forceinline__ float vertex_features(const ClothDev* __restrict__ cl, const int* __restrict__ vcloth,
                                                 const float4* __restrict__ x, const float4* __restrict__ p,
                                                 const float4* __restrict__ vel, float4* __restrict__ p_out,
                                                 float dt, const int* __restrict__ nbr,
                                                 const float* __restrict__ nbr_rest, int v, const float* s_mu,
                                                 const float* s_rsd, float gate, uint32_t* b,
                                                 float* __restrict__ feat_out) {
    // Loads in rounds (own state, neighbour ids and rest lengths, cloth id;
    // then the neighbour states and the cloth constants), so a thread waits
    // for a few memory round trips rather than one per neighbour.
    const int ci = vcloth[v];
    const float4 X = x[v];
    float4 P = ADVECT ? vel[v] : p[v];
    int nj[6];
    float nl[6];
#pragma unroll
    for (int k = 0; k < 6; k++) {
        nj[k] = nbr[(size_t)v * 6 + k];
        nl[k] = nbr_rest[(size_t)v * 6 + k];
    }
    const ClothDev& cd = cl[ci];
    const float h = cd.hscale;
    float4 nq[6], nv[6];
#pragma unroll
    for (int k = 0; k < 6; k++) {  // index 0 stands in for a missing neighbour
        if (ADVECT) {
            nq[k] = x[max(nj[k], 0)];
            nv[k] = vel[max(nj[k], 0)];
        } else {
            nq[k] = p[max(nj[k], 0)];
        }
    }
    if (ADVECT) {
        float4 vn;
        P = advect_one(cd, X, P, dt, &vn);
        p_out[v] = P;
#pragma unroll
        for (int k = 0; k < 6; k++) nq[k] = advect_one(cd, nq[k], nv[k], dt, &vn);
    }
    float f[22];
    f[0] = (P.x - X.x) / h;
    f[1] = (P.y - X.y) / h;
    f[2] = (P.z - X.z) / h;
    bool strained = !(gate > 0.0f);  // gate 0: every vertex; else some one-ring prim with |d - L| >= gate L
#pragma unroll
    for (int k = 0; k < 6; k++) {
        const int j = nj[k];
        float sx = 0.0f, sy = 0.0f, sz = 0.0f;
        if (j >= 0) {
            const float4 Q = nq[k];
            const float ex = Q.x - P.x, ey = Q.y - P.y, ez = Q.z - P.z;
            const float d = sqrtf(ex * ex + ey * ey + ez * ez);
            const float L = nl[k];
            strained |= !(fabsf(d - L) < gate * L);
            if (d > 0.0f) {
                const float sc = (d - L) / (L * d);
                sx = sc * ex;
                sy = sc * ey;
                sz = sc * ez;
            }
        }
        f[3 + 3 * k] = sx;
        f[4 + 3 * k] = sy;
        f[5 + 3 * k] = sz;
    }
    const bool pinned = P.w == 0.0f;
    f[21] = pinned ? 1.0f : 0.0f;
    if (feat_out)
        for (int k = 0; k < 22; k++) feat_out[(size_t)v * 22 + k] = f[k];
    // 32 BF16, two per word, in MFMA operand order
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
        for (int half = 0; half < 2; half++)
#pragma unroll
            for (int j = 0; j < 4; j += 2) {
                const int k0 = 16 * half + 4 * q + j;
                const uint32_t lo = k0 < 22 ? f2bf(norm_feature(f[k0], s_mu[k0], s_rsd[k0])) : 0u;
                const uint32_t hi = k0 + 1 < 22 ? f2bf(norm_feature(f[k0 + 1], s_mu[k0 + 1], s_rsd[k0 + 1])) : 0u;
                b[4 * q + 2 * half + j / 2] = lo | (hi << 16);
            }
    return (pinned || !strained) ? 0.0f : h;
}

__global__ __launch_bounds__(256) void k_ws_features(const ClothDev* __restrict__ cl, const int* __restrict__ vcloth,
                                                     const float4* __restrict__ x, const float4* __restrict__ p,
                                                     const int* __restrict__ nbr, const float* __restrict__ nbr_rest,
                                                     int nv, const float* __restrict__ norm, uint4* __restrict__ feat,
                                                     float* __restrict__ hs, float* __restrict__ feat_out, float gate) {
    __shared__ float s_mu[22], s_rsd[22];
    if (threadIdx.x < 22) {
        s_mu[threadIdx.x] = norm[threadIdx.x];
        s_rsd[threadIdx.x] = 1.0f / norm[22 + threadIdx.x];
    }
    __syncthreads();
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    uint32_t b[16];
    const float h = vertex_features<false>(cl, vcloth, x, p, nullptr, nullptr, 0.0f, nbr, nbr_rest, v, s_mu, s_rsd,
                                           gate, b, feat_out);
    uint4* row = feat + (size_t)v * 4;
    for (int q = 0; q < 4; q++) row[q] = make_uint4(b[4 * q], b[4 * q + 1], b[4 * q + 2], b[4 * q + 3]);
    hs[v] = h;
}

// Weights the network kernel keeps in LDS (the inference copies).
struct alignas(16) InfW {
    uint16_t i1[64 * 32];
    uint16_t i2[64 * 64];
    uint16_t i3[16 * 64];
    float b1[64], b2[64], b3[16];
};

static_assert(offsetof(MlpW, b3) - offsetof(MlpW, i1) == offsetof(InfW, b3) &&
                  sizeof(MlpW) - offsetof(MlpW, i1) == sizeof(InfW),
              "the inference copies must be one contiguous block laid out as InfW");

constexpr int MLP_WAVES = 8;

__device__ __forceinline__ bf16x4 as_bf16x4(uint32_t lo, uint32_t hi) {
    return __builtin_bit_cast(bf16x4, make_uint2(lo, hi));
}

__device__ __forceinline__ bf16x4 gelu4(f32x4 a) {
    float y0, y1, y2, y3;
    gelu(a[0], &y0, nullptr);
    gelu(a[1], &y1, nullptr);
    gelu(a[2], &y2, nullptr);
    gelu(a[3], &y3, nullptr);
    return pack_bf16(y0, y1, y2, y3);
}

// The network on 16 vertices, one wave: lane (r = l % 16, q = l / 16) passes
// the 16-byte feature chunk q of vertex r and gets back, in lanes q == 0,
// outputs 0..3 of vertex r (standardised).  Every layer is computed
// transposed, out^T = W * act^T, weights as the MFMA A operand and
// activations as B.  Lane (r, q) then holds, for output tile t, outputs
// 16t + 4q + i of vertex r -- exactly its B operand for K-step t of the next
// layer -- so activations never leave the registers; biases enter as the
// accumulator input.
__device__ __forceinline__ f32x4 mlp_group(const InfW& W, const uint4 cur, int r, int q) {
    const bf16x4 x0 = as_bf16x4(cur.x, cur.y), x1 = as_bf16x4(cur.z, cur.w);
    f32x4 acc[4];
    bf16x4 h1[4], h2[4];
    // layer 1: K = 32 (features); A operand of K-step st at 8q + 4st
    for (int t = 0; t < 4; t++) acc[t] = *(const f32x4*)(W.b1 + 16 * t + 4 * q);
    for (int st = 0; st < 2; st++) {
        const bf16x4 x = st == 0 ? x0 : x1;
        bf16x4 a[4];
        for (int t = 0; t < 4; t++) a[t] = *(const bf16x4*)(W.i1 + (16 * t + r) * 32 + 8 * q + 4 * st);
        for (int t = 0; t < 4; t++) acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[t], x, acc[t], 0, 0, 0);
    }
    for (int t = 0; t < 4; t++) h1[t] = gelu4(acc[t]);
    // layer 2: K = 64; A operand of K-step st at 16q + 4st
    for (int t = 0; t < 4; t++) acc[t] = *(const f32x4*)(W.b2 + 16 * t + 4 * q);
    for (int st = 0; st < 4; st++) {
        bf16x4 a[4];
        for (int t = 0; t < 4; t++) a[t] = *(const bf16x4*)(W.i2 + (16 * t + r) * 64 + 16 * q + 4 * st);
        for (int t = 0; t < 4; t++) acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[t], h1[st], acc[t], 0, 0, 0);
    }
    for (int t = 0; t < 4; t++) h2[t] = gelu4(acc[t]);
    // layer 3: 16 output rows (3 used), K = 64
    f32x4 o = *(const f32x4*)(W.b3 + 4 * q);
    for (int st = 0; st < 4; st++) {
        const bf16x4 a = *(const bf16x4*)(W.i3 + r * 64 + 16 * q + 4 * st);
        o = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, h2[st], o, 0, 0, 0);
    }
    return o;
}

__global__ __launch_bounds__(64 * MLP_WAVES, 4) void k_ws_mlp(const MlpW* __restrict__ gW, const uint4* __restrict__ feat,
                                                           const float* __restrict__ hs,
                                                           const float* __restrict__ norm, int nv,
                                                           float4* __restrict__ delta) {
    __shared__ InfW W;
    {
        const uint4* src = (const uint4*)gW->i1;  // i1, i2, i3, b1, b2, b3 are contiguous
        uint4* dst = (uint4*)&W;
        for (int i = threadIdx.x; i < (int)(sizeof(InfW) / 16); i += blockDim.x) dst[i] = src[i];
        __syncthreads();
    }
    const int wv = threadIdx.x >> 6, l = threadIdx.x & 63, r = l & 15, q = l >> 4;
    const float sdy0 = norm[44], sdy1 = norm[45], sdy2 = norm[46];
    const int ngroups = (nv + 15) / 16;
    const int stride = gridDim.x * MLP_WAVES;
    int g = blockIdx.x * MLP_WAVES + wv;
    auto fetch = [&](int grp) {
        const int v = grp * 16 + r;
        return (grp < ngroups && v < nv) ? feat[(size_t)v * 4 + q] : make_uint4(0u, 0u, 0u, 0u);
    };
    auto fetch_h = [&](int grp) {
        const int v = grp * 16 + r;
        return (q == 0 && grp < ngroups && v < nv) ? hs[v] : 0.0f;
    };
    uint4 next = fetch(g);
    float h_next = fetch_h(g);
    for (; g < ngroups; g += stride) {
        const uint4 cur = next;
        const float h = h_next;
        next = fetch(g + stride);
        h_next = fetch_h(g + stride);
        const int v = g * 16 + r;
        const f32x4 o = mlp_group(W, cur, r, q);
        // lane (r, 0) holds outputs 0..3 of vertex r
        if (q == 0 && v < nv) {
            float4 d = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (h != 0.0f) d = make_float4(h * (o[0] * sdy0), h * (o[1] * sdy1), h * (o[2] * sdy2), 0.0f);
            delta[v] = d;
        }
    }
}

// Features and network in one kernel (the solver's path, no raw features
// kept).  Each wave takes 64 vertices at a time: every lane computes one
// vertex's feature row and output scale into the wave's LDS area (chunk-major,
// so the 16-byte writes of consecutive lanes are consecutive), then the wave
// runs the network on the four groups of 16 from there.  Saves writing and
// re-reading 68 bytes per vertex against the two-kernel form.
struct alignas(16) InferStage {
    uint4 chunk[4][64];  // [q][vertex]: 16-byte feature chunk q of each vertex
    float hs[64];
};

template <bool ADVECT>
__global__ __launch_bounds__(64 * MLP_WAVES, 4) void k_ws_infer(
    const MlpW* __restrict__ gW, const ClothDev* __restrict__ cl, const int* __restrict__ vcloth,
    const float4* __restrict__ x, const float4* __restrict__ p, const float4* __restrict__ vel,
    float4* __restrict__ p_out, float dt, const int* __restrict__ nbr, const float* __restrict__ nbr_rest, int nv,
    const float* __restrict__ norm, float gate, float4* __restrict__ delta, unsigned long long* __restrict__ rs_zero,
    int n_rs) {
    __shared__ InfW W;
    __shared__ InferStage stage[MLP_WAVES];
    if (rs_zero)
        for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n_rs; i += gridDim.x * blockDim.x) rs_zero[i] = 0ull;
    __shared__ float s_mu[22], s_rsd[22];
    {
        const uint4* src = (const uint4*)gW->i1;  // i1, i2, i3, b1, b2, b3 are contiguous
        uint4* dst = (uint4*)&W;
        for (int i = threadIdx.x; i < (int)(sizeof(InfW) / 16); i += blockDim.x) dst[i] = src[i];
        if (threadIdx.x < 22) {
            s_mu[threadIdx.x] = norm[threadIdx.x];
            s_rsd[threadIdx.x] = 1.0f / norm[22 + threadIdx.x];
        }
        __syncthreads();
    }
    const int wv = threadIdx.x >> 6, l = threadIdx.x & 63, r = l & 15, q = l >> 4;
    InferStage& S = stage[wv];
    const float sdy0 = norm[44], sdy1 = norm[45], sdy2 = norm[46];
    const int nblk = (nv + 63) / 64;
    for (int g = blockIdx.x * MLP_WAVES + wv; g < nblk; g += gridDim.x * MLP_WAVES) {
        {
            const int v = g * 64 + l;
            uint32_t b[16];
            float h = 0.0f;
            if (v < nv) {
                h = vertex_features<ADVECT>(cl, vcloth, x, p, vel, p_out, dt, nbr, nbr_rest, v, s_mu, s_rsd, gate,
                                            b, nullptr);
            } else {
#pragma unroll
                for (int k = 0; k < 16; k++) b[k] = 0u;
            }
#pragma unroll
            for (int c = 0; c < 4; c++) S.chunk[c][l] = make_uint4(b[4 * c], b[4 * c + 1], b[4 * c + 2], b[4 * c + 3]);
            S.hs[l] = h;
        }
        // the wave's own LDS writes, visible to all its lanes
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (int s = 0; s < 4; s++) {
            const int vl = 16 * s + r, v = g * 64 + vl;
            const float h = S.hs[vl];
            // a group whose 16 output scales are all zero (pinned or quiet
            // vertices) predicts zero whatever the network says: skip it
            f32x4 o = {0.0f, 0.0f, 0.0f, 0.0f};
            if (__any(h != 0.0f)) o = mlp_group(W, S.chunk[q][vl], r, q);
            if (q == 0 && v < nv) {
                float4 d = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                if (h != 0.0f) d = make_float4(h * (o[0] * sdy0), h * (o[1] * sdy1), h * (o[2] * sdy2), 0.0f);
                delta[v] = d;
            }
        }
        // every lane has read the stage before the next group overwrites it
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// Residual of every distance prim with and without the predicted delta, per
// cloth: the max (FP64 bit patterns, r >= 0) and the sum of squares in
// 2^-40 fixed point (r^2 clamped to 1; integer sums do not depend on the
// order of the atomics).  rs = [max0 | max1 | sq0 | sq1], ncl entries each.
// The prim list is in cloth order, so each thread keeps running values over
// RESID_PER_THREAD prims (stride 256) and flushes them with atomics only when
// its cloth changes or at the end; a wave whose lanes all end on one cloth
// reduces first.
constexpr int RESID_PER_THREAD = 64;  // long runs per thread: few flushes (measured 8: 0.57 ms, 64: 0.30 ms)

template <int RULE>
__device__ __forceinline__ void flush_resid(unsigned long long* rs, int ncl, int c, double a, double b,
                                            unsigned long long sa, unsigned long long sb) {
    if (RULE == 0) {
        atomicMax(&rs[c], (unsigned long long)__double_as_longlong(a));
        atomicMax(&rs[ncl + c], (unsigned long long)__double_as_longlong(b));
    } else {
        atomicAdd(&rs[2 * ncl + c], sa);
        atomicAdd(&rs[3 * ncl + c], sb);
    }
}

template <int RULE>  // 0: max only, 1: squared sums only (what the fallback rule reads)
__global__ __launch_bounds__(256) void k_ws_resid2(const int4* __restrict__ dist, int n, const float4* __restrict__ p,
                                                   const float4* __restrict__ delta,
                                                   unsigned long long* __restrict__ rs, int ncl) {
    const long long base = (long long)blockIdx.x * (256 * RESID_PER_THREAD) + threadIdx.x;
    int cl = -1;
    double a = 0.0, b = 0.0;
    unsigned long long sa = 0, sb = 0;
    // prims in batches of RESID_BATCH: records first, then their four
    // gathers each, so a batch waits for two memory round trips
    constexpr int RESID_BATCH = 2;
    for (int k0 = 0; k0 < RESID_PER_THREAD; k0 += RESID_BATCH) {
        int4 qb[RESID_BATCH];
        float4 P0b[RESID_BATCH], P1b[RESID_BATCH], D0b[RESID_BATCH], D1b[RESID_BATCH];
#pragma unroll
        for (int u = 0; u < RESID_BATCH; u++) {
            const long long j = base + 256ll * (k0 + u);
            qb[u] = j < n ? dist[j] : make_int4(0, 0, 0, -1);
        }
#pragma unroll
        for (int u = 0; u < RESID_BATCH; u++) {
            P0b[u] = p[qb[u].x];
            P1b[u] = p[qb[u].y];
            D0b[u] = delta[qb[u].x];
            D1b[u] = delta[qb[u].y];
        }
        bool done = false;
#pragma unroll
        for (int u = 0; u < RESID_BATCH; u++) {
            if (base + 256ll * (k0 + u) >= n) {
                done = true;
                break;
            }
            const int4 q = qb[u];
            if (q.w != cl) {
                if (cl >= 0) flush_resid<RULE>(rs, ncl, cl, a, b, sa, sb);
                cl = q.w;
                a = b = 0.0;
                sa = sb = 0;
            }
            float4 P0 = P0b[u], P1 = P1b[u], D0 = D0b[u], D1 = D1b[u];
            float L = __int_as_float(q.z);
            const float r0 = rel_res(P0, P1, L);
            float4 Q0 = make_float4(P0.x + D0.x, P0.y + D0.y, P0.z + D0.z, 0.0f);
            float4 Q1 = make_float4(P1.x + D1.x, P1.y + D1.y, P1.z + D1.z, 0.0f);
            const float r1 = rel_res(Q0, Q1, L);
            if (RULE == 0) {
                a = fmax(a, (double)r0);
                b = fmax(b, (double)r1);
            } else {
                sa += sq_fixed(r0);
                sb += sq_fixed(r1);
            }
        }
        if (done) break;
    }
    int c0 = __shfl(cl, 0);
    unsigned long long same = __ballot(cl == c0 || cl < 0);
    if (same == ~0ull) {
        for (int off = 32; off > 0; off >>= 1) {
            if (RULE == 0) {
                a = fmax(a, __shfl_xor(a, off));
                b = fmax(b, __shfl_xor(b, off));
            } else {
                sa += __shfl_xor(sa, off);
                sb += __shfl_xor(sb, off);
            }
        }
        if ((threadIdx.x & 63) == 0 && c0 >= 0) flush_resid<RULE>(rs, ncl, c0, a, b, sa, sb);
    } else if (cl >= 0) {
        flush_resid<RULE>(rs, ncl, cl, a, b, sa, sb);
    }
}

// Fallback rule and per-cloth sweep budget of the substep: keep the
// prediction iff its residual (rule 0: max, rule 1: RMS) does not exceed the
// unwarmed start's; a kept cloth runs the warm budget, a discarded one the
// plain solver's full budget.
__global__ __launch_bounds__(256) void k_ws_decide(const unsigned long long* __restrict__ rs, int ncl, int rule,
                                                   int warm_sweeps, int full_sweeps, int* __restrict__ cap,
                                                   int* __restrict__ keep) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncl) return;
    const bool k = ws_keep(rs, ncl, c, rule);
    keep[c] = k ? 1 : 0;
    cap[c] = k ? warm_sweeps : full_sweeps;
}

__global__ __launch_bounds__(256) void k_ws_apply(const int* __restrict__ vcloth, float4* __restrict__ p,
                                                  const float4* __restrict__ delta, const int* __restrict__ keep,
                                                  int nv) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    if (!keep[vcloth[v]]) return;
    float4 P = p[v], D = delta[v];
    p[v] = make_float4(P.x + D.x, P.y + D.y, P.z + D.z, P.w);
}

__global__ __launch_bounds__(256) void k_ws_normalize(float* __restrict__ X, float* __restrict__ Y, long long n,
                                                      const float* __restrict__ norm) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int k = 0; k < 22; k++) X[i * 22 + k] = norm_feature(X[i * 22 + k], norm[k], 1.0f / norm[22 + k]);
    for (int k = 0; k < 3; k++) Y[i * 3 + k] = Y[i * 3 + k] / norm[44 + k];
}

// Gauss-Seidel correction target: y = (p_after - p_before) / h.
__global__ __launch_bounds__(256) void k_ws_target(const ClothDev* __restrict__ cl, const int* __restrict__ vcloth,
                                                   const float4* __restrict__ pa, const float4* __restrict__ pb,
                                                   int nv, float* __restrict__ y) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    float h = cl[vcloth[v]].hscale;
    float4 A = pa[v], B = pb[v];
    y[(size_t)v * 3 + 0] = (A.x - B.x) / h;
    y[(size_t)v * 3 + 1] = (A.y - B.y) / h;
    y[(size_t)v * 3 + 2] = (A.z - B.z) / h;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static inline int nb256(long long n) { return (int)((n + 255) / 256); }

int mlp_nparam() { return NPARAM; }
size_t mlp_wbytes() { return sizeof(MlpW); }

void mlp_pack(hipStream_t s, const float* prm, void* W) {
    hipLaunchKernelGGL(k_mlp_pack, dim3(16), dim3(256), 0, s, prm, (MlpW*)W);
    PBD_HIP(hipGetLastError());
}

void mlp_forward(hipStream_t s, const void* W, const float* X, int n, float* Y) {
    if (n <= 0) return;
    int blocks = (n + 16 * FWD_WAVES - 1) / (16 * FWD_WAVES);
    hipLaunchKernelGGL(k_mlp_forward, dim3(blocks), dim3(64 * FWD_WAVES), 0, s, (const MlpW*)W, X, n, Y);
    PBD_HIP(hipGetLastError());
}

// Gradient of the mean loss over samples idx[0..n) into grad[NPARAM]
// (grad[NPARAM] = loss).  partial must hold ceil(n/16) * (NPARAM+1) floats.
void mlp_grad(hipStream_t s, const void* W, const float* X, const float* T, const int* idx, int n, float inv_total,
              float* partial, float* grad) {
    int waves = (n + 15) / 16;
    int blocks = (waves + TRAIN_WAVES - 1) / TRAIN_WAVES;
    PBD_HIP(hipMemsetAsync(partial, 0, sizeof(float) * (size_t)blocks * TRAIN_WAVES * (NPARAM + 1), s));
    hipLaunchKernelGGL(k_mlp_train, dim3(blocks), dim3(64 * TRAIN_WAVES), 0, s, (const MlpW*)W, X, T, idx, n,
                       inv_total, partial);
    hipLaunchKernelGGL(k_mlp_reduce, dim3((NPARAM + 1 + 15) / 16), dim3(256), 0, s, partial, waves, grad);
    PBD_HIP(hipGetLastError());
}

void mlp_adam(hipStream_t s, float* w, const float* g, float* m, float* v, float lr, float b1, float b2, float eps,
              int step) {
    float c1 = (float)(1.0 - std::pow((double)b1, step)), c2 = (float)(1.0 - std::pow((double)b2, step));
    hipLaunchKernelGGL(k_adam, dim3(nb256(NPARAM)), dim3(256), 0, s, w, g, m, v, NPARAM, lr, b1, b2, eps, c1, c2);
    PBD_HIP(hipGetLastError());
}

// BF16 feature rows of the network kernel: 64 bytes per vertex.
size_t ws_feature_bytes(int nv) { return (size_t)std::max(nv, 1) * 64; }

// Workgroups of `kernel` (at `threads`, no dynamic LDS) the current device
// holds at once: CUs x resident workgroups per CU, queried once per device
// and kernel.
int resident_blocks(const void* kernel, int threads) {
    static std::mutex m;
    static std::map<std::pair<int, const void*>, int> cache;
    int dev = 0;
    PBD_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(m);
    auto it = cache.find({dev, kernel});
    if (it != cache.end()) return it->second;
    int ncu = 0, per_cu = 0;
    PBD_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    PBD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0));
    const int n = std::max(1, ncu) * std::max(1, per_cu);
    cache[{dev, kernel}] = n;
    return n;
}

void ws_infer(hipStream_t s, const void* W, const ClothDev* cl, const int* vcloth, const float4* x, float4* p,
              const float4* vel, float dt, const int* nbr, const float* nbr_rest, int nv, float4* delta,
              float* feat_out, const float* norm, void* feat_scratch, float* hs_scratch, float gate,
              unsigned long long* rs_zero, int n_rs) {
    PBD_CHECK(!(vel && feat_out), PBD_ERR_STATE, "internal: advect-on-load inference keeps no raw features");
    PBD_CHECK(!(rs_zero && feat_out), PBD_ERR_STATE, "internal: only the fused inference clears the sums");
    if (nv <= 0) return;
    if (!feat_out) {
        const int max_blocks_fused = resident_blocks((const void*)k_ws_infer<true>, 64 * MLP_WAVES);
        const int nblk = (nv + 63) / 64;
        const int blocks = std::min((nblk + MLP_WAVES - 1) / MLP_WAVES, max_blocks_fused);
        if (vel)
            hipLaunchKernelGGL(k_ws_infer<true>, dim3(blocks), dim3(64 * MLP_WAVES), 0, s, (const MlpW*)W, cl, vcloth,
                               x, (const float4*)nullptr, vel, p, dt, nbr, nbr_rest, nv, norm, gate, delta, rs_zero,
                               n_rs);
        else
            hipLaunchKernelGGL(k_ws_infer<false>, dim3(blocks), dim3(64 * MLP_WAVES), 0, s, (const MlpW*)W, cl,
                               vcloth, x, (const float4*)p, (const float4*)nullptr, (float4*)nullptr, 0.0f, nbr,
                               nbr_rest, nv, norm, gate, delta, rs_zero, n_rs);
        PBD_HIP(hipGetLastError());
        return;
    }
    uint4* feat = (uint4*)feat_scratch;
    hipLaunchKernelGGL(k_ws_features, dim3(nb256(nv)), dim3(256), 0, s, cl, vcloth, x, p, nbr, nbr_rest, nv, norm,
                       feat, hs_scratch, feat_out, gate);
    PBD_HIP(hipGetLastError());
    int groups = (nv + 15) / 16;
    int blocks = std::min((groups + MLP_WAVES - 1) / MLP_WAVES, resident_blocks((const void*)k_ws_mlp, 64 * MLP_WAVES));
    hipLaunchKernelGGL(k_ws_mlp, dim3(blocks), dim3(64 * MLP_WAVES), 0, s, (const MlpW*)W, (const uint4*)feat,
                       (const float*)hs_scratch, norm, nv, delta);
    PBD_HIP(hipGetLastError());
}


void ws_resid2(hipStream_t s, const int4* dist, int n, const float4* p, const float4* delta, unsigned long long* rs,
               int ncl, int rule) {
    if (n <= 0) return;
    int blocks = (int)(((long long)n + 256 * RESID_PER_THREAD - 1) / (256 * RESID_PER_THREAD));
    if (rule == 1)
        hipLaunchKernelGGL(k_ws_resid2<1>, dim3(blocks), dim3(256), 0, s, dist, n, p, delta, rs, ncl);
    else
        hipLaunchKernelGGL(k_ws_resid2<0>, dim3(blocks), dim3(256), 0, s, dist, n, p, delta, rs, ncl);
    PBD_HIP(hipGetLastError());
}

void ws_decide(hipStream_t s, const unsigned long long* rs, int ncl, int rule, int warm_sweeps, int full_sweeps,
               int* cap, int* keep) {
    if (ncl <= 0) return;
    hipLaunchKernelGGL(k_ws_decide, dim3(nb256(ncl)), dim3(256), 0, s, rs, ncl, rule, warm_sweeps, full_sweeps, cap,
                       keep);
    PBD_HIP(hipGetLastError());
}

void ws_apply(hipStream_t s, const int* vcloth, float4* p, const float4* delta, const int* keep, int nv) {
    if (nv <= 0) return;
    hipLaunchKernelGGL(k_ws_apply, dim3(nb256(nv)), dim3(256), 0, s, vcloth, p, delta, keep, nv);
    PBD_HIP(hipGetLastError());
}

void ws_normalize(hipStream_t s, float* X, float* Y, long long n, const float* norm) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_ws_normalize, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, X, Y, n, norm);
    PBD_HIP(hipGetLastError());
}



void ws_target(hipStream_t s, const ClothDev* cl, const int* vcloth, const float4* pa, const float4* pb, int nv,
               float* y) {
    if (nv <= 0) return;
    hipLaunchKernelGGL(k_ws_target, dim3(nb256(nv)), dim3(256), 0, s, cl, vcloth, pa, pb, nv, y);
    PBD_HIP(hipGetLastError());
}

}  // namespace pbd
