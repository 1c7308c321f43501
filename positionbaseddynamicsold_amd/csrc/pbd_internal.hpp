// Internal declarations shared by the host model (model.cpp), the C ABI
// (capi.cpp) and the HIP kernels (*.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/pbd.h"

namespace pbd {

// Error carried from anywhere in the library to the C ABI boundary.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define PBD_HIP(call)                                                                    \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess)                                                            \
            throw ::pbd::Error(PBD_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define PBD_CHECK(cond, code, msg)                       \
    do {                                                 \
        if (!(cond)) throw ::pbd::Error((code), (msg));  \
    } while (0)

// Owning device buffer.
template <class T>
struct DBuf {
    T* ptr = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { release(); }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        n = 0;
    }
    void alloc(size_t count) {
        if (count == n && ptr) return;
        release();
        if (count) PBD_HIP(hipMalloc(&ptr, count * sizeof(T)));
        n = count;
    }
    void upload(const std::vector<T>& h, hipStream_t s) {
        alloc(h.size());
        if (!h.empty()) PBD_HIP(hipMemcpyAsync(ptr, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void download(std::vector<T>& h, hipStream_t s) const {
        h.resize(n);
        if (n) PBD_HIP(hipMemcpyAsync(h.data(), ptr, n * sizeof(T), hipMemcpyDeviceToHost, s));
    }
};

// ---------------------------------------------------------------------------
// Device-side records
// ---------------------------------------------------------------------------

// Per-cloth constants read by the kernels.
struct ClothDev {
    float gdt[3];        // gravity * dt
    float damp;          // powf(damping, dt)
    float wind[3];
    float kd;            // wind_drag * dt
    float wind_drag;
    int sphere_on;
    float sphere_c[3];
    float sphere_r;
    int box_on;
    float box_lo[3];
    float box_hi[3];
    int vbase, nv;       // vertex range in the packed arrays
    int ktab;            // offset of this cloth's per-sweep stiffness tables
    int whole;           // 1 if the cloth is a single LDS tile
    int tile_begin, tile_end;  // its tiles' slots in the current plan
    float hscale;        // mean rest length (warm-start feature scale)
    double tol;          // residual tolerance of this cloth (< 0: the solver's)
};

// AdvectGPUKernel (PBD_cuda.cu:949-979): v += g*dt*m ; v *= damping^dt ;
// p = x + v*dt, with m -> w; optional wind drag v += ((wind - v)*(drag*dt))*w.
__device__ __forceinline__ float4 advect_one(const ClothDev& c, float4 X, float4 V, float dt, float4* Vout) {
    float w = X.w;
    float vx = V.x + c.gdt[0] * w, vy = V.y + c.gdt[1] * w, vz = V.z + c.gdt[2] * w;
    if (c.wind_drag != 0.0f) {
        vx = vx + ((c.wind[0] - vx) * c.kd) * w;
        vy = vy + ((c.wind[1] - vy) * c.kd) * w;
        vz = vz + ((c.wind[2] - vz) * c.kd) * w;
    }
    vx = vx * c.damp;
    vy = vy * c.damp;
    vz = vz * c.damp;
    *Vout = make_float4(vx, vy, vz, 0.0f);
    return make_float4(X.x + vx * dt, X.y + vy * dt, X.z + vz * dt, w);
}

// One LDS tile: a set of loaded vertices (interior first) and the prims whose
// vertices are all loaded, grouped by colour.
struct TileDev {
    int cloth;
    int whole;           // 1: the tile is the whole cloth (runs complete substeps)
    int load_off;        // offset into tile_verts
    int n_load;
    int n_int;           // interior vertices = first n_int of the loaded list
    int dist_off;        // offset into tile_dist of colour 0
    int bend_off;        // offset into tile_bend of colour 0
    int own_off;         // offset into tile_own (distance prims owned by the tile)
    int n_own;
    int col_off;         // offset into tile_cols: (ncolors+1) dist offsets, (ncolors+1) bend offsets, then for
                         // partial tiles plan_sweeps x ncolors dist counts and as many bend counts
    int ncolors;
    int plan_sweeps;     // partial tiles: sweeps per launch the cone was planned for; sweep s of an n-sweep
                         // launch runs the first counts[(s + plan_sweeps - n) * ncolors + c] prims of colour c
                         // (0: whole tiles, every prim of the colour)
};

// Kernel-wide parameters of one solver launch.
struct SolveArgs {
    float dt;
    int sweep_base;      // global sweep index of this launch's first sweep
    int nsweeps;         // sweeps this launch for multi-tile cloths' tiles
    int whole_nsweeps;   // sweeps this launch for single-tile cloths' tiles
    int check_every;
    double tol;
    int kmax;            // length of each stiffness table
    int fused;           // whole tiles: 1 = advect on load + integrate on store
    int chunk;           // chunk index (selects the residual slot)
    int no_coll;         // 1 = skip the per-sweep analytic colliders (applied by the caller)
    int copy_spent;      // 1 = a multi-tile cloth past its sweep budget is copied forward
    // warm start folded into the load (first launch of a substep only):
    // p + delta for the vertices of cloths with keep[cloth] != 0
    const float4* delta;
    const int* keep;
    int advect_load;     // 1: multi-tile cloths' tiles advect x, v on load (first launch only)
    int integrate_end;   // 1: a tile storing its cloth's last sweep also integrates (x, v); no tile of
                         //    a launch where this can happen reads x
    int coherent_load;   // 1: p_in was written earlier in this launch by other workgroups (a queued
                         //    later chunk): loaded past every cache
    int queue_ts;        // queued launches: sweeps per chunk; a tile whose cloth has a further chunk
                         //    stores p_out write-through for it (0: not queued)
};

// ---------------------------------------------------------------------------
// Host model of one cloth
// ---------------------------------------------------------------------------
struct Cloth {
    // input
    int rx = 0, ry = 0;                   // grid dims (0 = general mesh)
    std::vector<float> pos0;              // 3*nv
    std::vector<int> tris;                // 3*nt
    std::vector<int> user_pairs;          // explicit distance prims
    bool has_user_pairs = false;
    std::vector<int> user_quads;          // explicit bending prims (e0, e1, o0, o1)
    bool has_user_quads = false;
    std::vector<float> user_rest;         // explicit rest values, prim order (empty: from pos0)
    std::vector<float> w_in;              // explicit inverse-mass weights (empty: 1; pins 0 either way)
    std::vector<float> x_start, v_start, p_start;  // start state (empty: pos0, 0, x)
    std::vector<int> pins;
    unsigned ct = PBD_DISTANCE;
    float stiff[3] = {1.0f, 1.0f, 1.0f};
    pbd_cloth_params prm{};
    // built
    int nv = 0, nt = 0, ne = 0, nb = 0;
    std::vector<int> prim_list;           // 2*np (start,count)
    std::vector<int> indices;
    std::vector<float> rest;
    std::vector<float> w;
    std::vector<int> adj_start_num;       // 2*np
    std::vector<int> adj;
    std::vector<int> color;
    std::vector<int> sorted;
    std::vector<int> worksets;            // 2*ncolors
    int ncolors = 0;
    int color_rounds = 0;
    int vbase = 0;                        // offset in packed vertex arrays
    int pbase = 0;                        // offset in packed prim arrays
    int tile_begin = 0, tile_end = 0;     // range in the tile table
    std::vector<int> gd_cut, gb_cut;      // per colour: [begin, end) in the batch colour-major prim arrays
    int rd_lo = 0, rd_hi = 0;             // range in the cloth-ordered residual list
    bool whole = false;
    int max_loaded = 0;
    // tile plan sums for the bytes a projection moves: loaded, interior and
    // owned-prim counts over the cloth's tiles, the prim-record bytes of one
    // launch of n sweeps (index n; whole tiles: of one sweep, index 1) and
    // the vertices such a launch loads (index n; whole tiles: all, index 1)
    long long tiles_loaded = 0, tiles_interior = 0, tiles_owned = 0;
    std::vector<double> tile_record_bytes, tile_loaded_by_sweeps;
    int last_sweeps = 0;
    double tol = -1.0;                    // per-cloth residual tolerance (< 0: the solver's)
    int np() const { return ne + nb; }
};

// Host-side helpers shared by model.cpp and the tests of the C ABI.
uint64_t splitmix64(uint64_t* s);
float uniform_pm1(uint64_t* s);
float distance_powf(const float* a, const float* b);
void make_grid(const pbd_grid_desc& d, std::vector<float>& pos, std::vector<int>& tris);
void grid_distance_pairs(int rx, int ry, std::vector<int>& pairs);
void mesh_edge_pairs(const std::vector<int>& tris, std::vector<int>& pairs);
void bending_quads(const std::vector<int>& tris, std::vector<int>& quads);
float atan2_poly(float y, float x);
float dihedral_angle(const float* e0, const float* e1, const float* o0, const float* o1);
void build_prims(Cloth& c);
void build_adjacency(Cloth& c);
void read_txt_cloth(const char* topol, const char* constr, std::vector<float>& pos,
                    std::vector<int>& tris, std::vector<int>& pairs, bool& has_pairs);
// Cache text format (cache_io.cpp).  x, v, p are 3*nv host arrays of the
// cloth's current state; precision is the significant digits written (0: 6).
void write_cache(const Cloth& c, const float* x, const float* v, const float* p, const char* mesh_path,
                 const char* constr_path, int precision);
void read_cache(const char* mesh_path, const char* constr_path, Cloth& c);
// Edges of a triangle list in first-sighting order (triangle by triangle,
// edges (v0,v1), (v1,v2), (v2,v0)), endpoints (lower, higher); tedge[3t+k]
// is the id of triangle t's edge k.
void build_edge_table(const int* tris, int nt, std::vector<int>& tedge, std::vector<int2>& edges);

// Broadphase grid: dense (reference) or hashed table.
struct HashCfg {
    int mode;            // 0 dense, 1 hashed
    float start[3];      // centre - size*cell/2 (dense) / origin of cell coordinates (hashed)
    float cell[3];
    int size[3];
    int table_bits;
    int ncells;          // dense: size product; hashed: 1 << table_bits
};

// Broadphase + CCD working set (device buffers reused between calls).
struct Broadphase {
    HashCfg cfg{};
    int nt = 0;
    long long npairs = 0, ncontacts = 0;
    DBuf<uint32_t> keys, ids, k2, v2, pv, pf, t1, t2, uv, uf;
    DBuf<int> raw, hist, hist2, cstart, cend, cnt, off, scan_tmp, flag, pos, cv, cf, vcnt;
    DBuf<float> ctmp_t, ctmp_n, ct, cn;
    DBuf<double> ctmp_w, cw;
    DBuf<unsigned long long> acc;
    DBuf<int4> tcell;    // centroid cell coordinates per triangle
    DBuf<float4> sw_box; // swept-box broadphase: face boxes (lo, hi)
    DBuf<int4> sw_cell;  // and their cell ranges (lo, hi)
    DBuf<int> sw_err, sw_vin;
    // running counts (host values the drivers read anyway): detections,
    // triangles, cell entries, raw pairs, distinct pairs, contacts
    long long st_dets = 0, st_tris = 0, st_entries = 0, st_raw = 0, st_pairs = 0, st_contacts = 0, st_resolved = 0;
    // edge-edge: per-triangle edge ids, edge endpoints (lower, higher), the
    // distinct candidate edge pairs and the contacts (a0, a1, b0, b1)
    DBuf<int> tedge;
    DBuf<int2> edges;
    DBuf<int> voff, vnbr;  // vertex -> mesh neighbours (CSR) of the edge table
    int nedges = 0;
    DBuf<uint32_t> eua, eub;
    DBuf<int4> ecq;
    DBuf<float> ect, ecn;
    DBuf<double> ecw;
    long long ee_npairs = 0, ee_ncontacts = 0;
    long long st_ee_raw = 0, st_ee_pairs = 0, st_ee_contacts = 0;
    void build(hipStream_t s, const HashCfg& g, const float4* p, const int* tris, int nt);
    long long pairs(hipStream_t s, const float4* p, const int* tris, int nv);
    long long unique_pairs(hipStream_t s, int total, int nt, int nv);
    long long unique_into(hipStream_t s, int total, int nt, int nv, DBuf<uint32_t>& oa, DBuf<uint32_t>& ob);
    void set_edges(hipStream_t s, const std::vector<int>& tedge, const std::vector<int2>& edges);
    // after build(): edge-edge candidates and contacts (contacts resolved with
    // the vertex-face ones by resolve())
    long long ee_pairs(hipStream_t s, const float4* p, const int* tris);
    long long ee_narrow(hipStream_t s, const float4* x, const float4* p, float th);
    // Swept-box candidate set (replaces build + pairs); x = start of substep.
    long long pairs_swept(hipStream_t s, const HashCfg& g, const float4* x, const float4* p, const int* tris, int nt,
                          int nv, float th);
    long long narrow(hipStream_t s, const float4* x, const float4* p, const int* tris, float th);
    // min_n2: faces with |(p2 - p1) x (p3 - p1)|^2 <= min_n2 are skipped
    // (VFResolveTest's degenerate-face test, 1e-6 in the reference)
    void resolve(hipStream_t s, const float4* x, float4* p, const int* tris, int nv, float th, int passes,
                 float min_n2);
};
void device_scan(hipStream_t s, const int* in, int* out, int n, int* tmp);
// Broadphase build into caller device arrays with no allocation: the sort's
// scratch is a static device array of sh_noalloc_scratch_bytes() per GPU.
constexpr int SH_NOALLOC_MAX_TRIS = 1 << 20;
size_t sh_noalloc_scratch_bytes();
void sh_build_noalloc(hipStream_t s, int device, const HashCfg& g, const float* pos, int nv, const int* tris, int nt,
                      int* sorted_keys, int* sorted_ids, int* raw_keys, int* cell_start, int* cell_end);

// Warm-start MLP (mlp.hip).
int mlp_nparam();
size_t mlp_wbytes();
void mlp_pack(hipStream_t s, const float* prm, void* W);
void mlp_forward(hipStream_t s, const void* W, const float* X, int n, float* Y);
void mlp_grad(hipStream_t s, const void* W, const float* X, const float* T, const int* idx, int n, float inv_total,
              float* partial, float* grad);
void mlp_adam(hipStream_t s, float* w, const float* g, float* m, float* v, float lr, float b1, float b2, float eps,
              int step);
// Predicted corrections.  With vel, p is written: the advected positions of
// x and vel (advect on load, no raw features); otherwise p is read.
// rs_zero (fused form only): n_rs words the kernel clears on the way (the
// fallback's residual sums, so that no separate fill runs before them).
void ws_infer(hipStream_t s, const void* W, const ClothDev* cl, const int* vcloth, const float4* x, float4* p,
              const float4* vel, float dt, const int* nbr, const float* nbr_rest, int nv, float4* delta,
              float* feat_out, const float* norm, void* feat_scratch, float* hs_scratch, float gate,
              unsigned long long* rs_zero = nullptr, int n_rs = 0);
// Per-prim relative residual |d - L| / L in FP32 (it only decides whether the
// prediction is kept; the reductions over it are exact).
__device__ __forceinline__ float rel_res(float4 a, float4 b, float L) {
    float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
    float d = sqrtf(dx * dx + dy * dy + dz * dz);
    return fabsf(d - L) / L;
}
// r^2 (clamped to 1) in 2^-40 fixed point: sums of these do not depend on the
// order in which they are added.
__device__ __forceinline__ unsigned long long sq_fixed(float r) {
    const float q = fminf(r * r, 1.0f);
    return (unsigned long long)((double)q * 1099511627776.0);  // 2^40
}
// The fallback rule on one cloth's residual sums rs = [max0 | max1 | sq0 | sq1]
// (ncl entries each): keep the prediction iff its residual does not exceed
// the unwarmed start's (rule 0: max |C|/L, bit patterns of non-negative
// doubles; rule 1: sums of squares in fixed point).
__device__ __forceinline__ bool ws_keep(const unsigned long long* rs, int ncl, int c, int rule) {
    if (rule == 1) return rs[3 * ncl + c] <= rs[2 * ncl + c];
    const double a = __longlong_as_double((long long)rs[c]), b = __longlong_as_double((long long)rs[ncl + c]);
    return b <= a;
}
// The keep decision folded into the tile queue's schedule kernel (it also
// clears the sweep counters and residual slots and records each cloth's
// final p buffer, the work of separate small kernels otherwise).
struct QueueDecide {
    const unsigned long long* rs;
    int rule, warm_sweeps, full_sweeps;
    int* keep;
    int* cap;       // the budgets the decision writes (the projection's cap)
    bool in_launch; // the queued tile launch computes the residuals and decides (rs unused)
};
size_t ws_feature_bytes(int nv);
void ws_normalize(hipStream_t s, float* X, float* Y, long long n, const float* norm);
void ws_resid2(hipStream_t s, const int4* dist, int n, const float4* p, const float4* delta, unsigned long long* rs,
               int ncl, int rule);
void ws_decide(hipStream_t s, const unsigned long long* rs, int ncl, int rule, int warm_sweeps, int full_sweeps,
               int* cap, int* keep);
void ws_apply(hipStream_t s, const int* vcloth, float4* p, const float4* delta, const int* keep, int nv);
void ws_target(hipStream_t s, const ClothDev* cl, const int* vcloth, const float4* pa, const float4* pb, int nv,
               float* y);

// Tile plan for the whole batch.
struct TilePlan {
    std::vector<TileDev> tiles;
    std::vector<int> verts;               // global vertex ids
    std::vector<uint32_t> dist;           // packed (i0 | i1 << 16), local indices
    std::vector<float> dist_rest;
    std::vector<uint32_t> bend;           // 2 words per prim: (i0|i1<<16), (i2|i3<<16)
    std::vector<float> bend_rest;
    std::vector<uint32_t> own;            // owned distance prims (local pair)
    std::vector<float> own_rest;
    std::vector<int> cols;                // per tile colour offsets
    std::vector<int> partial_order;       // launch order of the partial tiles only
    std::vector<int> whole_order;         // launch order of the whole tiles only
    // neighbours of each partial tile (CSR over tiles): the tiles of its
    // cloth whose interior it loads or which load its interior, itself
    // included -- the tiles a later chunk of it must wait for
    std::vector<int> nbr_off, nbr;
    // plan statistics: loaded and interior vertices, loaded distance prims,
    // and distance prim updates of one launch of `sweeps` sweeps
    long long stat_loaded = 0, stat_interior = 0, stat_dist_loaded = 0, stat_dist_phase_needed = 0;
};
// split_whole: a cloth that fits one tile still gets a (single, halo-free)
// multi-launch tile instead of a whole tile.
void plan_tiles(std::vector<Cloth>& cloths, int max_load, int sweeps_per_launch, bool split_whole, TilePlan& plan);

// ---------------------------------------------------------------------------
// Kernel launchers (implemented in the .hip files)
// ---------------------------------------------------------------------------
// One colouring round over all np prims; the prims left uncoloured are
// counted into *uncolored.
int launch_color_round(hipStream_t s, const int* prim_npts, const int* adj_sn, const int* adj, int* color,
                       int* prd, int* uncolored, int np);
// Ints of `hist` that radix_sort_pairs needs for n keys: the per-block digit
// counts (256 x ceil(n / 4096)), their exclusive scan, and the block sums of
// that scan.
inline size_t radix_hist_ints(long long n) {
    const size_t h = (size_t)256 * (size_t)((std::max(1ll, n) + 4095) / 4096);
    return 2 * h + (h + 1023) / 1024 + 2;
}
void radix_sort_pairs(hipStream_t s, uint32_t* k, uint32_t* v, uint32_t* k2, uint32_t* v2, int n,
                      int bits, int* hist);
void launch_advect(hipStream_t s, const ClothDev* cloths, const int* vcloth, const float4* x,
                   float4* v, float4* p, int nv_begin, int nv_end, float dt, int skip_whole);
void launch_integrate(hipStream_t s, const ClothDev* cloths, const int* vcloth, float4* x, float4* v,
                      float4* p0, float4* p1, const int* pfinal, int pend, int nv_begin, int nv_end, float dt,
                      int skip_whole);
void launch_consolidate(hipStream_t s, const int* vcloth, float4* p0, float4* p1, const int* pfinal, int pend,
                        int n);
void launch_final_buffers(hipStream_t s, const ClothDev* cl, int ncl, int start, int launched, int ts,
                          int iterations, const int* cap, int* pfinal);
void launch_project_colour(hipStream_t s, const int4* dist, int d_begin, int d_end,
                           const int4* bend_idx, const float2* bend_rk, int b_begin, int b_end,
                           const float* ktab_d, const float* ktab_b, int kmax, int sweep, const int* cap,
                           float4* p);
void launch_colliders(hipStream_t s, const ClothDev* cloths, const int* vcloth, float4* p,
                      int nv_begin, int nv_end, const int* cap = nullptr, int sweep = 0);
// Incidence lists of the Jacobi path (see k_jacobi in solver.hip).
constexpr int JAC_SLOTS = 8;
struct JacobiDev {
    int nv;
    const int* ddeg;     // distance incidences per vertex
    const int2* dslot;   // first JAC_SLOTS of them, slot-major: dslot[k * nv + i]
    const int* doff;     // nv + 1 offsets of the rest into dovf
    const int2* dovf;
    const int* boff;     // nv + 1 offsets into bent / bmeta
    const int4* bent;    // bending prim vertices (global ids)
    const int2* bmeta;   // (rest bits, role)
};
void launch_jacobi(hipStream_t s, const JacobiDev& J, const ClothDev* cloths, const int* vcloth, const float* ktab_d,
                   const float* ktab_b, int kmax, int sweep, const int* cap, float omega, const float4* pin,
                   float4* pout, int nv_begin, int nv_end);
void launch_residual(hipStream_t s, const int4* dist, int n, const float4* p,
                     unsigned long long* rmax, double* rsum, int* rcnt);
void launch_tiles(hipStream_t s, int nblocks, const int* order, const TileDev* tiles,
                  const int* verts, const uint32_t* dist, const float* dist_rest,
                  const uint32_t* bend, const float* bend_rest, const uint32_t* own,
                  const float* own_rest, const int* cols, const ClothDev* cloths,
                  const float* ktab_d, const float* ktab_b, float4* x, float4* v,
                  const float4* p_in, float4* p_out, unsigned long long* res_slots,
                  int nslots_cloth, int* sweeps_out, const int* sweep_cap, const SolveArgs& a,
                  int threads, size_t lds_bytes, bool with_bend, const int* n_items = nullptr);
// One launch for every chunk of a substep under per-cloth budgets: the
// workgroups take (tile, chunk) items from a queue in which every item comes
// after the items it depends on (its neighbour tiles' previous chunk).
struct TileQueue {
    const int* item_tile;   // items: tile slot and chunk
    const int* item_chunk;
    const int* n_items;     // device count
    int* head;              // next item to take (0 at the start)
    int* done;              // per tile: last chunk finished (-1 at the start)
    const int* nbr_off;     // tile neighbours (CSR), the tile itself included
    const int* nbr;
    float4* pbuf0;          // chunk k reads pbuf[start ^ (k & 1)], writes the other
    float4* pbuf1;
    int start;
    int ts;                 // sweeps per chunk
    int iterations;
    int* timeouts;          // waits that gave up (bounded spins; 0 when every wait ended normally)
    // The warm start's keep decision made inside the launch (launch_tiles_decide).
    // Every tile's first chunk is a fixed item (stat_item: the single-tile
    // cloths' tiles, then the multi-tile cloths' tiles cloth by cloth); the
    // later chunks of a cloth whose budget outlasts the first are appended
    // (dyn_tile / dyn_chunk) by the tile that completes its cloth's decision.
    // Everything below except the lists is zero when the launch starts.
    const int* stat_item;
    int n_static;
    int n_multi;            // multi-tile cloths: each is decided once
    int* dyn_tile;
    int* dyn_chunk;
    int* ctr;               // [next fixed item, next appended item, appended reserved, appended published,
                            //  multi-tile cloths decided]
    int* posted;            // per cloth: tiles whose residual sums are in
    int* decided;           // per cloth c at [c * QC_DECIDED_STRIDE]: 0 until decided, then 1 + keep
    int* finished;          // per tile: chunks finished (chunk c done: c + 1)
    unsigned long long* rs; // per cloth: [max0 | max1 | sq0 | sq1] as ws_keep reads them
    int rule, warm_sweeps, full_sweeps;
    int* keep;              // written per cloth by the decision, with cap, pfinal and a zeroed sweeps
    int* cap;
    int* pfinal;
    int* launches;          // incremented once per deciding launch
};
// Counter slots of TileQueue::ctr, one 128-byte line each (the workgroups'
// polls of one do not stall the atomics on another), and the spacing of the
// per-cloth decision flags for the same reason.
enum {
    QC_STAT = 0,
    QC_DYN_HEAD = 32,
    QC_DYN_TAIL = 64,
    QC_DYN_READY = 96,
    QC_DECIDED = 128,
    QC_ABORT = 160,
    QC_COUNT = 192,
    QC_DECIDED_STRIDE = 32
};
// The queued launch with the keep decision inside it: a tile's first chunk
// loads the unwarmed start of its cone, adds its owned distance prims' relative
// residuals (with and without the prediction) to its cloth's sums, waits
// until every tile of the cloth has done so, applies the fallback rule, and
// goes on from the kept prediction or from the unwarmed start.  A cloth's
// tiles wait for each other, so the launch needs every multi-tile cloth's
// tiles resident at once (at most nblocks - 1 tiles per cloth).
void launch_tiles_decide(hipStream_t s, int nblocks, const TileDev* tiles, const int* verts, const uint32_t* dist,
                         const float* dist_rest, const uint32_t* bend, const float* bend_rest, const uint32_t* own,
                         const float* own_rest, const int* cols, const ClothDev* cloths, const float* ktab_d,
                         const float* ktab_b, float4* x, float4* v, unsigned long long* res_slots, int ncl,
                         int* sweeps_out, const int* sweep_cap, const SolveArgs& a, const TileQueue& q, int threads,
                         size_t lds_bytes, bool with_bend);
void launch_tiles_queue(hipStream_t s, int nblocks, const TileDev* tiles, const int* verts, const uint32_t* dist,
                        const float* dist_rest, const uint32_t* bend, const float* bend_rest, const uint32_t* own,
                        const float* own_rest, const int* cols, const ClothDev* cloths, const float* ktab_d,
                        const float* ktab_b, float4* x, float4* v, unsigned long long* res_slots, int ncl,
                        int* sweeps_out, const int* sweep_cap, const SolveArgs& a, const TileQueue& q, int threads,
                        size_t lds_bytes, bool with_bend);
// Builds the queue of a substep on the device (one workgroup): resets head
// and every tile's done, splits the first-chunk order into the tiles of
// cloths whose budget outlasts the first chunk (the chain) and the others,
// and lays the items out as [chain, chunk 0][others, part 0][chain, chunk 1]
// [others, part 1] ... so that the chain's chunks run beside the others'
// work.  chain_tmp / bulk_tmp: scratch of n_order ints each.
void launch_queue_schedule(hipStream_t s, const int* order, int n_order, const TileDev* tiles, int n_tiles,
                           const int* cap, int ts, int nchunks, int* chain_tmp, int* bulk_tmp, int* item_tile,
                           int* item_chunk, int* n_items, int* head, int* done);
// The same with the keep decision (writing cap) and the substep's small
// bookkeeping in the one workgroup: sweeps[ncl] and res_slots[n_slots]
// cleared, pfinal[c] = the p buffer each cloth ends in.
void launch_queue_schedule_decide(hipStream_t s, const QueueDecide& d, const ClothDev* cl, int ncl, int* cap,
                                  int* sweeps, unsigned long long* res_slots, int n_slots, int* pfinal, int start,
                                  int iterations, const int* order, int n_order, const TileDev* tiles, int n_tiles,
                                  int ts, int nchunks, int* chain_tmp, int* bulk_tmp, int* item_tile, int* item_chunk,
                                  int* n_items, int* head, int* done);
// Compacts the multi-tile order (layout order[j * 8 + g]; order_cloth: each
// entry's cloth, -1 for padding) to the tiles whose cloth's budget cap
// exceeds min_cap, rebalanced over the 8 groups, length into *n_out.
void launch_follow_list(hipStream_t s, const int* order, const int* order_cloth, int n_order, const int* cap,
                        int min_cap, int* out, int* n_out);
int tiles_per_cu(bool with_bend, int threads, size_t lds_bytes);
// Opts the tile kernels in to `bytes` of dynamic LDS on the current device;
// returns what they may use (bytes, or 64 KiB if the runtime refuses).
size_t tiles_lds_optin(size_t bytes);
constexpr int TILE_STATIC_LDS = 256;  // bytes reserved for the tile kernel's static LDS (block reductions)
void launch_fill_u64(hipStream_t s, unsigned long long* p, size_t n, unsigned long long v);
void launch_fill_int(hipStream_t s, int* p, size_t n, int v);
void launch_min_sweeps(hipStream_t s, int* sweeps, const int* cap, int sw, int n);
void launch_copy(hipStream_t s, const float4* a, float4* b, size_t n);

}  // namespace pbd
