// extern "C" boundary of libpbd.so: context, model build, init pipeline and
// solver orchestration.  Exceptions never cross this file's functions.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>

#include "pbd_internal.hpp"

using namespace pbd;

static thread_local std::string g_err;

struct pbd_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<Cloth> cloths;
    bool inited = false;
    pbd_solver_params sp{};
    int lds_max = 65536;
    // state
    DBuf<float4> x, v, pbuf[2];
    int pcur = 0;
    // after a tile projection each cloth's result sits in p buffer pfinal[c]
    // until integrate (or a consolidation) gathers it into pbuf[pcur]
    DBuf<int> pfinal;
    bool pfinal_valid = false;
    DBuf<int> vcloth;
    std::vector<ClothDev> cl_h;
    DBuf<ClothDev> cl;
    int nv_total = 0;
    // global-path prims, batch colour-major
    DBuf<int4> gdist, gbidx, rdist;
    DBuf<float2> gbrc;
    std::vector<int> gd_off, gb_off;
    int ncolors_max = 0;
    // Jacobi path: vertex -> incident prims over the packed vertices, built
    // on first use (layout: JacobiDev)
    DBuf<int> jddeg, jdoff, jboff;
    DBuf<int2> jdslot, jdovf, jbmeta;
    DBuf<int4> jbent;
    JacobiDev jac{};
    bool jac_built = false;
    // stiffness tables
    DBuf<float> ktab_d, ktab_b;
    int kmax = 0;
    // tiles
    TilePlan plan;
    int plan_sweeps = 0, plan_max_load = 0, plan_split = 0;
    DBuf<TileDev> tiles;
    // launch orders: single-tile cloths' tiles, multi-tile cloths' tiles, and
    // both (the first launch of a substep runs them side by side)
    DBuf<int> tverts, tcols, order_whole, order_partial, order_first, tile_ids;
    // later chunks under per-cloth budgets: the multi-tile tiles still
    // needed, compacted on the device, and the list's length
    DBuf<int> order_follow, follow_len, order_partial_cloth;
    // the one-launch form of the same (TileQueue): tile neighbours, per-tile
    // done chunks, the queue's items and scratch, and the wait timeouts
    DBuf<int> tnbr_off, tnbr, tdone, q_tile, q_chunk, q_tmp, q_count, q_timeouts;
    // the queued launch that makes the keep decision itself: its fixed
    // first-chunk items, the block it needs zeroed at its start (residual
    // sums, per-cloth and per-tile counters), its multi-tile cloths and the
    // most tiles one of them has
    DBuf<int> q_stat;
    DBuf<unsigned long long> q_zero;
    int q_n_multi = 0, q_tiles_max = 0;
    int num_cu = 0;
    // workgroups the tile kernel keeps resident on the whole GPU, for the
    // (kernel, threads, LDS) it was last asked for
    int walk_key_threads = 0, walk_key_bend = -1, walk_slots = 0;
    size_t walk_key_lds = 0;
    DBuf<uint32_t> tdist, tbend, town;
    DBuf<float> tdist_rest, tbend_rest, town_rest;
    bool any_whole = false, any_partial = false;
    // residual / sweep bookkeeping
    DBuf<unsigned long long> res_slots, rmax;
    DBuf<double> rsum;
    DBuf<int> sweeps;
    int nslots = 0;
    // standalone broadphase / CCD (parity-test entry points)
    Broadphase sbp;
    DBuf<float4> sx, sp4;
    DBuf<int> stris;
    int snv = 0, snt = 0;
    uint64_t scene_hash = 0;  // of the standalone scene and its detection settings (scene_fingerprint)
    bool sh_built = false, ccd_built = false;
    // in-solver collision
    pbd_collision_params coll{};
    HashCfg ccfg{};
    Broadphase cbp;
    DBuf<int> gtris;
    int gnt = 0;
    // learned warm start
    struct WarmStart {
        bool params_ready = false, enabled = false, nbr_built = false;
        DBuf<float> prm, m, v, grad, partial, X, Y, norm;
        std::vector<float> norm_h;
        bool samples_normalized = false;
        DBuf<unsigned char> W;
        int adam_step = 0;
        DBuf<int> nbr, perm;
        DBuf<float> nbr_rest;
        long long ns = 0;
        DBuf<float4> delta, psave;
        DBuf<unsigned long long> rs;  // per cloth: max residual without/with, squared sums without/with
        DBuf<int> keep;               // per cloth: 1 if the last prediction was kept
        std::vector<int> used;
        bool used_pending = false;  // r0/r1 hold a decision not yet read back
        DBuf<int> cap;              // per-cloth sweep budget of the current substep
        DBuf<uint4> feat;           // BF16 feature rows of the network kernel
        DBuf<float> hs;             // per-vertex output scale (0 = pinned)
    } ws;
    bool sweeps_pending = false;    // the device sweep counts are newer than Cloth::last_sweeps
    // substep captured as a graph, one per starting p buffer; the p buffer
    // the replay ends in is graph_end
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    int graph_end[2] = {0, 0};
    bool graph_pvalid[2] = {false, false};  // pfinal_valid at the end of the captured substep
    // device-side state snapshot
    DBuf<float4> snap_x, snap_v;
    bool snap_valid = false;
    // profiling
    bool prof_on = false;
    pbd_profile prof{};
    std::vector<hipEvent_t> ev_pool;
    int ev_used = 0;
    ~pbd_ctx() {
        for (auto e : ev_pool) (void)hipEventDestroy(e);
        for (auto& g : graph)
            if (g) (void)hipGraphExecDestroy(g);
    }
};

namespace {

template <class F>
int guard(F&& f) {
    try {
        f();
        return PBD_OK;
    } catch (const Error& e) {
        g_err = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_err = "host allocation failed";
        return PBD_ERR_LIMIT;
    } catch (const std::exception& e) {
        g_err = e.what();
        return PBD_ERR_ARG;
    }
}

Cloth& cloth_of(pbd_ctx* c, int id) {
    PBD_CHECK(c, PBD_ERR_ARG, "null context");
    PBD_CHECK(id >= 0 && id < (int)c->cloths.size(), PBD_ERR_ARG, "cloth id out of range");
    return c->cloths[id];
}

void set_device(pbd_ctx* c) { PBD_HIP(hipSetDevice(c->device)); }

// Captured substeps bake in buffer pointers and by-value kernel arguments;
// every change of those drops them.
void drop_graphs(pbd_ctx* c) {
    for (auto& g : c->graph) {
        if (g) (void)hipGraphExecDestroy(g);
        g = nullptr;
    }
}

pbd_cloth_params default_params() {
    pbd_cloth_params p{};
    p.damping = 1.0f;
    p.gravity[1] = -9.8f;
    return p;
}

// Per-cloth kernel constants for substep length dt.
void fill_cloth_dev(pbd_ctx* c) {
    drop_graphs(c);
    float dt = c->sp.dt;
    c->cl_h.resize(c->cloths.size());
    for (size_t i = 0; i < c->cloths.size(); i++) {
        const Cloth& k = c->cloths[i];
        ClothDev& d = c->cl_h[i];
        d = ClothDev{};
        for (int a = 0; a < 3; a++) {
            d.gdt[a] = k.prm.gravity[a] * dt;
            d.wind[a] = k.prm.wind[a];
            d.sphere_c[a] = k.prm.sphere_c[a];
            d.box_lo[a] = k.prm.box_lo[a];
            d.box_hi[a] = k.prm.box_hi[a];
        }
        d.damp = powf(k.prm.damping, dt);
        d.kd = k.prm.wind_drag * dt;
        d.wind_drag = k.prm.wind_drag;
        d.sphere_on = k.prm.sphere_on;
        d.sphere_r = k.prm.sphere_r;
        d.box_on = k.prm.box_on;
        d.vbase = k.vbase;
        d.nv = k.nv;
        d.ktab = (int)i * c->kmax;
        d.whole = k.whole ? 1 : 0;
        d.tile_begin = k.tile_begin;
        d.tile_end = k.tile_end;
        double sum = 0.0;
        int ne = 0;
        for (int p = 0; p < k.np(); p++)
            if (k.prim_list[2 * p + 1] == 2) {
                sum += k.rest[p];
                ne++;
            }
        d.hscale = ne ? (float)(sum / ne) : 1.0f;
        d.tol = k.tol;
    }
    c->cl.upload(c->cl_h, c->stream);
}

// Stiffness per sweep: distance prims follow the reference kernel rule
// k = 1 - powf(1 - s, 1/(iter+1)) (PBD_cuda.cu:1282); bending prims the
// constant k' = 1 - (1 - s)^(1/I) of Mueller et al. 2007 with I = iterations.
void fill_ktabs(pbd_ctx* c) {
    int I = std::max(1, c->sp.iterations);
    c->kmax = I;
    std::vector<float> kd((size_t)c->cloths.size() * I), kb((size_t)c->cloths.size() * I);
    for (size_t i = 0; i < c->cloths.size(); i++) {
        float sd = c->cloths[i].stiff[0], sb = c->cloths[i].stiff[1];
        float kbb = 1 - powf(1 - sb, 1.0 / I);
        for (int s = 0; s < I; s++) {
            kd[i * I + s] = 1 - powf(1 - sd, 1.0 / (s + 1));
            kb[i * I + s] = kbb;
        }
    }
    c->ktab_d.upload(kd, c->stream);
    c->ktab_b.upload(kb, c->stream);
}

int effective_tile_sweeps(const pbd_ctx* c) {
    if (c->coll.enabled) return 1;  // contacts are resolved after every sweep
    int s = c->sp.tile_sweeps > 0 ? c->sp.tile_sweeps : c->sp.check_every;
    return std::max(1, s);
}

int effective_max_load(const pbd_ctx* c) {
    int lim = (c->lds_max - TILE_STATIC_LDS) / 16;
    if (c->sp.tile_max_verts > 0) lim = std::min(lim, c->sp.tile_max_verts);
    return std::max(64, lim);
}

void upload_plan(pbd_ctx* c) {
    TilePlan& p = c->plan;
    c->tiles.upload(p.tiles, c->stream);
    c->tverts.upload(p.verts, c->stream);
    c->tcols.upload(p.cols, c->stream);
    c->tdist.upload(p.dist, c->stream);
    c->tdist_rest.upload(p.dist_rest, c->stream);
    c->tbend.upload(p.bend, c->stream);
    c->tbend_rest.upload(p.bend_rest, c->stream);
    c->town.upload(p.own, c->stream);
    c->town_rest.upload(p.own_rest, c->stream);
    c->order_whole.upload(p.whole_order, c->stream);
    c->order_partial.upload(p.partial_order, c->stream);
    std::vector<int> first(p.whole_order);  // single-tile cloths first: their blocks run longest
    first.insert(first.end(), p.partial_order.begin(), p.partial_order.end());
    c->order_first.upload(first, c->stream);
    std::vector<int> pcl(p.partial_order.size());
    for (size_t i = 0; i < pcl.size(); i++) pcl[i] = p.partial_order[i] >= 0 ? p.tiles[p.partial_order[i]].cloth : -1;
    c->order_partial_cloth.upload(pcl, c->stream);
    c->order_follow.alloc(std::max<size_t>(1, p.partial_order.size()));
    c->follow_len.alloc(1);
    c->tnbr_off.upload(p.nbr_off, c->stream);
    c->tnbr.upload(p.nbr.empty() ? std::vector<int>(1, 0) : p.nbr, c->stream);
    c->tdone.alloc(std::max<size_t>(1, p.tiles.size()));
    c->q_tmp.alloc(2 * std::max<size_t>(1, first.size()));
    c->q_count.alloc(2);  // item count, head
    if (!c->q_timeouts.n) {
        c->q_timeouts.alloc(2);  // wait timeouts, deciding launches
        PBD_HIP(hipMemsetAsync(c->q_timeouts.ptr, 0, 2 * sizeof(int), c->stream));
    }
    std::vector<int> ids(p.tiles.size());
    for (size_t t = 0; t < ids.size(); t++) ids[t] = (int)t;
    c->tile_ids.upload(ids, c->stream);
    {
        // first chunks of the deciding queue: single-tile cloths' tiles (their
        // blocks run longest), then the multi-tile cloths' tiles cloth by cloth
        std::vector<int> st;
        st.reserve(p.tiles.size());
        for (size_t t = 0; t < p.tiles.size(); t++)
            if (p.tiles[t].whole) st.push_back((int)t);
        for (size_t t = 0; t < p.tiles.size(); t++)
            if (!p.tiles[t].whole) st.push_back((int)t);
        c->q_stat.upload(st.empty() ? std::vector<int>(1, 0) : st, c->stream);
        c->q_n_multi = 0;
        c->q_tiles_max = 0;
        for (const Cloth& k : c->cloths)
            if (!k.whole && k.tile_end > k.tile_begin) {
                c->q_n_multi++;
                c->q_tiles_max = std::max(c->q_tiles_max, k.tile_end - k.tile_begin);
            }
        // [rs: 4 ncl words, padded to 128 bytes][counters][decision flags][posted][finished]
        const size_t ncl = c->cloths.size();
        const size_t ints = QC_COUNT + QC_DECIDED_STRIDE * ncl + ncl + p.tiles.size();
        c->q_zero.alloc((4 * ncl + 15) / 16 * 16 + (ints + 1) / 2);
    }
    drop_graphs(c);
    c->any_whole = false;
    c->any_partial = false;
    for (auto& t : p.tiles) (t.whole ? c->any_whole : c->any_partial) = true;
}

// Items the substep's tile queue can hold: every tile of the first chunk and
// every multi-tile tile in each later chunk.  Allocated outside graph
// capture (allocation is not allowed inside it).
void ensure_queue(pbd_ctx* c) {
    const int ts = std::max(1, effective_tile_sweeps(c));
    const size_t nchunks = (size_t)(std::max(1, c->sp.iterations) + ts - 1) / ts;
    const size_t need = std::max<size_t>(1, c->order_first.n + nchunks * c->order_partial.n);
    if (c->q_tile.n < need) {
        drop_graphs(c);
        c->q_tile.alloc(need);
        c->q_chunk.alloc(need);
    }
}

// Grid of the list-walking tile launches: every resident workgroup slot of
// the GPU once, rounded up to a multiple of 8 (the XCD layout of the lists).
int walk_grid(pbd_ctx* c) {
    const bool bend = !c->plan.bend.empty();
    const int threads = c->sp.block_threads > 0 ? c->sp.block_threads : 1024;
    const size_t lds = (size_t)std::max(1, c->plan_max_load) * sizeof(float4);
    if (threads != c->walk_key_threads || lds != c->walk_key_lds || (int)bend != c->walk_key_bend) {
        c->walk_slots = std::max(1, c->num_cu) * tiles_per_cu(bend, threads, lds);
        c->walk_key_threads = threads;
        c->walk_key_lds = lds;
        c->walk_key_bend = (int)bend;
    }
    return (c->walk_slots + 7) / 8 * 8;
}

void replan_if_needed(pbd_ctx* c) {
    int s = effective_tile_sweeps(c), ml = effective_max_load(c), sp = c->sp.split_whole ? 1 : 0;
    if (s == c->plan_sweeps && ml == c->plan_max_load && sp == c->plan_split) return;
    // a failed plan leaves no valid one behind: the next call plans again
    c->plan_sweeps = 0;
    c->plan_max_load = 0;
    plan_tiles(c->cloths, ml, s, sp != 0, c->plan);
    c->plan_sweeps = s;
    c->plan_max_load = ml;
    c->plan_split = sp;
    upload_plan(c);
    fill_cloth_dev(c);
}

void ensure_slots(pbd_ctx* c, int nchunks) {
    int need = std::max(2, nchunks + 1);
    if (need > c->nslots) {
        drop_graphs(c);
        c->res_slots.alloc((size_t)need * c->cloths.size());
        c->nslots = need;
    }
}

// ---- model init -----------------------------------------------------------
void do_init(pbd_ctx* c, int color_rounds) {
    PBD_CHECK(!c->cloths.empty(), PBD_ERR_STATE, "no cloth created");
    PBD_CHECK(color_rounds > 0, PBD_ERR_ARG, "color_rounds must be positive");
    set_device(c);
    hipStream_t s = c->stream;
    // constraints + adjacency per cloth, packed offsets
    int vbase = 0, pbase = 0;
    long long adj_total = 0;
    for (auto& k : c->cloths) {
        build_prims(k);
        build_adjacency(k);
        k.vbase = vbase;
        k.pbase = pbase;
        vbase += k.nv;
        pbase += k.np();
        adj_total += (long long)k.adj.size();
    }
    PBD_CHECK(adj_total < (1ll << 31), PBD_ERR_LIMIT, "adjacency exceeds 2^31 entries");
    c->nv_total = vbase;
    int np = pbase;
    // ---- GPU colouring (EdgeColoringGPU, PBD_cuda.cu:487-544)
    std::vector<int> npts(np), adj_sn((size_t)2 * np), adj;
    adj.reserve((size_t)adj_total);
    for (auto& k : c->cloths) {
        // cloth-local CSR rebased to global prim ids and offsets
        int base = (int)adj.size();
        for (int p = 0; p < k.np(); p++) {
            npts[k.pbase + p] = k.prim_list[2 * p + 1];
            adj_sn[2 * (size_t)(k.pbase + p)] = base + k.adj_start_num[2 * p];
            adj_sn[2 * (size_t)(k.pbase + p) + 1] = k.adj_start_num[2 * p + 1];
        }
        for (int q : k.adj) adj.push_back(q + k.pbase);
    }
    std::vector<int> color;
    int rounds = 0;
    if (np > 0) {
        DBuf<int> d_npts, d_sn, d_adj, d_color, d_prd, d_unc;
        d_npts.upload(npts, s);
        d_sn.upload(adj_sn, s);
        d_adj.upload(adj, s);
        std::vector<int> minus1(np, -1);
        d_color.upload(minus1, s);
        d_prd.upload(minus1, s);
        d_unc.alloc(1);
        int unc = np;
        while (unc > 0) {
            PBD_CHECK(rounds < color_rounds, PBD_ERR_COLOR, "colouring did not converge in the allowed rounds");
            PBD_HIP(hipMemsetAsync(d_unc.ptr, 0, sizeof(int), s));
            launch_color_round(s, d_npts.ptr, d_sn.ptr, d_adj.ptr, d_color.ptr, d_prd.ptr, d_unc.ptr, np);
            PBD_HIP(hipMemcpyAsync(&unc, d_unc.ptr, sizeof(int), hipMemcpyDeviceToHost, s));
            PBD_HIP(hipStreamSynchronize(s));
            rounds++;
        }
        // ---- device stable colour sort (SortEdgesColors, PBD_cuda.cu:546-587)
        std::vector<uint32_t> ids(np);
        for (int i = 0; i < np; i++) ids[i] = (uint32_t)i;
        d_color.download(color, s);
        PBD_HIP(hipStreamSynchronize(s));
        int cmax = 0;
        for (int q : color) cmax = std::max(cmax, q + 1);
        int bits = 1;
        while ((1 << bits) < cmax) bits++;
        DBuf<uint32_t> kk, vv, k2, v2;
        DBuf<int> hist;
        kk.upload(std::vector<uint32_t>(color.begin(), color.end()), s);
        vv.upload(ids, s);
        k2.alloc(np);
        v2.alloc(np);
        hist.alloc(radix_hist_ints(np));
        radix_sort_pairs(s, kk.ptr, vv.ptr, k2.ptr, v2.ptr, np, bits, hist.ptr);
        std::vector<uint32_t> sorted_ids;
        vv.download(sorted_ids, s);
        PBD_HIP(hipStreamSynchronize(s));
        c->ncolors_max = cmax;
        // per-cloth colour data
        for (auto& k : c->cloths) {
            k.color.assign(color.begin() + k.pbase, color.begin() + k.pbase + k.np());
            k.sorted.clear();
            k.color_rounds = rounds;
        }
        for (uint32_t g : sorted_ids) {
            // prims of one cloth are contiguous in global id
            auto it = std::upper_bound(c->cloths.begin(), c->cloths.end(), (int)g,
                                       [](int val, const Cloth& k) { return val < k.pbase; });
            Cloth& k = *(it - 1);
            k.sorted.push_back((int)g - k.pbase);
        }
        for (auto& k : c->cloths) {
            k.ncolors = 0;
            for (int q : k.color) k.ncolors = std::max(k.ncolors, q + 1);
            k.worksets.assign((size_t)2 * k.ncolors, 0);
            for (int q : k.color) k.worksets[2 * q + 1]++;
            int run = 0;
            for (int q = 0; q < k.ncolors; q++) {
                k.worksets[2 * q] = run;
                run += k.worksets[2 * q + 1];
            }
        }
    } else {
        for (auto& k : c->cloths) {
            k.color.clear();
            k.sorted.clear();
            k.worksets.clear();
            k.ncolors = 0;
        }
    }
    // ---- batch colour-major prim arrays (global path) and cloth-ordered residual list
    std::vector<int4> gd, rd, gb;
    std::vector<float2> gbr;
    c->gd_off.assign(c->ncolors_max + 1, 0);
    c->gb_off.assign(c->ncolors_max + 1, 0);
    for (auto& k : c->cloths) {
        k.gd_cut.assign(2 * (size_t)k.ncolors, 0);
        k.gb_cut.assign(2 * (size_t)k.ncolors, 0);
    }
    for (int col = 0; col < c->ncolors_max; col++) {
        c->gd_off[col] = (int)gd.size();
        c->gb_off[col] = (int)gb.size();
        for (size_t ci = 0; ci < c->cloths.size(); ci++) {
            Cloth& k = c->cloths[ci];
            if (col >= k.ncolors) continue;
            k.gd_cut[2 * col] = (int)gd.size();
            k.gb_cut[2 * col] = (int)gb.size();
            int st = k.worksets[2 * col], n = k.worksets[2 * col + 1];
            for (int j = st; j < st + n; j++) {
                int p = k.sorted[j];
                const int* ix = &k.indices[k.prim_list[2 * p]];
                if (k.prim_list[2 * p + 1] == 2) {
                    float r = k.rest[p];
                    int rb;
                    memcpy(&rb, &r, 4);
                    gd.push_back(make_int4(ix[0] + k.vbase, ix[1] + k.vbase, rb, (int)ci));
                } else {
                    gb.push_back(make_int4(ix[0] + k.vbase, ix[1] + k.vbase, ix[2] + k.vbase, ix[3] + k.vbase));
                    float cb;
                    int ci_i = (int)ci;
                    memcpy(&cb, &ci_i, 4);
                    gbr.push_back(make_float2(k.rest[p], cb));
                }
            }
            k.gd_cut[2 * col + 1] = (int)gd.size();
            k.gb_cut[2 * col + 1] = (int)gb.size();
        }
    }
    c->gd_off[c->ncolors_max] = (int)gd.size();
    c->gb_off[c->ncolors_max] = (int)gb.size();
    for (size_t ci = 0; ci < c->cloths.size(); ci++) {
        Cloth& k = c->cloths[ci];
        k.rd_lo = (int)rd.size();
        for (int p = 0; p < k.np(); p++) {
            if (k.prim_list[2 * p + 1] != 2) continue;
            const int* ix = &k.indices[k.prim_list[2 * p]];
            float r = k.rest[p];
            int rb;
            memcpy(&rb, &r, 4);
            rd.push_back(make_int4(ix[0] + k.vbase, ix[1] + k.vbase, rb, (int)ci));
        }
        k.rd_hi = (int)rd.size();
    }
    c->gdist.upload(gd, s);
    c->gbidx.upload(gb, s);
    c->gbrc.upload(gbr, s);
    c->rdist.upload(rd, s);
    // ---- state buffers (InitGPUBuffers, PBD_cuda.cu:707-750)
    // a cloth continued from a cache starts from its cached x, v and p
    std::vector<float4> X(c->nv_total), V(c->nv_total, make_float4(0, 0, 0, 0)), P;
    std::vector<int> vc(c->nv_total);
    bool any_p = false;
    for (Cloth& k : c->cloths) any_p = any_p || !k.p_start.empty();
    if (any_p) P.resize(c->nv_total);
    for (size_t ci = 0; ci < c->cloths.size(); ci++) {
        Cloth& k = c->cloths[ci];
        const float* xs = k.x_start.empty() ? k.pos0.data() : k.x_start.data();
        for (int i = 0; i < k.nv; i++) {
            X[k.vbase + i] = make_float4(xs[3 * i], xs[3 * i + 1], xs[3 * i + 2], k.w[i]);
            if (!k.v_start.empty())
                V[k.vbase + i] = make_float4(k.v_start[3 * i], k.v_start[3 * i + 1], k.v_start[3 * i + 2], 0.0f);
            if (any_p)
                P[k.vbase + i] = k.p_start.empty() ? X[k.vbase + i]
                                                   : make_float4(k.p_start[3 * i], k.p_start[3 * i + 1],
                                                                 k.p_start[3 * i + 2], k.w[i]);
            vc[k.vbase + i] = (int)ci;
        }
    }
    c->x.upload(X, s);
    c->v.upload(V, s);
    c->pbuf[0].upload(any_p ? P : X, s);
    c->pbuf[1].upload(any_p ? P : X, s);
    c->pcur = 0;
    c->vcloth.upload(vc, s);
    c->rmax.alloc(c->cloths.size());
    c->rsum.alloc(c->cloths.size());
    c->sweeps.alloc(c->cloths.size());
    PBD_HIP(hipMemsetAsync(c->sweeps.ptr, 0, sizeof(int) * c->cloths.size(), s));
    c->pfinal.alloc(c->cloths.size());
    c->pfinal_valid = false;
    fill_ktabs(c);
    c->plan_sweeps = 0;
    c->plan_max_load = 0;
    replan_if_needed(c);
    PBD_HIP(hipStreamSynchronize(s));
    c->inited = true;
}

// ---- profiling -------------------------------------------------------------
hipEvent_t next_event(pbd_ctx* c) {
    if (c->ev_used == (int)c->ev_pool.size()) {
        hipEvent_t e;
        PBD_HIP(hipEventCreate(&e));
        c->ev_pool.push_back(e);
    }
    return c->ev_pool[c->ev_used++];
}

// Sums the durations of the recorded (start, stop) pairs; call after a sync.
void collect_events(pbd_ctx* c) {
    if (c->ev_used > 0) PBD_HIP(hipEventSynchronize(c->ev_pool[c->ev_used - 1]));
    for (int i = 0; i + 1 < c->ev_used; i += 2) {
        float ms = 0.0f;
        PBD_HIP(hipEventElapsedTime(&ms, c->ev_pool[i], c->ev_pool[i + 1]));
        c->prof.tile_ms += ms;
        c->prof.tile_launches++;
    }
    c->ev_used = 0;
}

// Algorithmic bytes per prim per sweep and per vertex (fused advect + integrate).
constexpr double kDistBytes = 68.0, kBendBytes = 132.0, kVertBytes = 100.0;  // per prim per sweep; per vertex
constexpr double kAdvectBytes = 52.0, kIntegrateBytes = 48.0;                // the two parts of kVertBytes

// ---- broadphase configuration ----------------------------------------------
// Grid start as SpatialHashSystem::InitSH computes it (SpatialHashSystem.cpp:88-98):
// shift = (gridSize * cellSize) / 2 in float, start = centre - shift.
HashCfg make_hash_cfg(const pbd_hash_desc& d) {
    PBD_CHECK(d.mode == 0 || d.mode == 1, PBD_ERR_ARG, "hash mode must be 0 (dense) or 1 (hashed)");
    HashCfg g{};
    g.mode = d.mode;
    for (int k = 0; k < 3; k++) {
        PBD_CHECK(d.cell[k] > 0.0f, PBD_ERR_ARG, "cell size must be positive");
        PBD_CHECK(d.mode == 1 || d.size[k] > 0, PBD_ERR_ARG, "grid size must be positive");
        float shift = (float)((unsigned)d.size[k]) * d.cell[k] / 2;
        g.start[k] = d.centre[k] - shift;
        g.cell[k] = d.cell[k];
        g.size[k] = d.size[k];
    }
    g.table_bits = d.table_bits;
    if (d.mode == 0) {
        long long nc = (long long)d.size[0] * d.size[1] * d.size[2];
        PBD_CHECK(nc < (1ll << 30), PBD_ERR_LIMIT, "dense grid larger than 2^30 cells");
        g.ncells = (int)nc;
    } else {
        PBD_CHECK(d.table_bits >= 1 && d.table_bits <= 28, PBD_ERR_ARG, "table_bits must be in [1, 28]");
        g.ncells = 1 << d.table_bits;
    }
    return g;
}

// FNV-1a over 8-byte words of a standalone scene and the settings of its
// detection (x, p, triangles, sizes, thickness, grid description): tells
// whether an edge-edge detection runs on the scene and settings of the
// previous vertex-face one.
uint64_t scene_fingerprint(const float* x, const float* p, int nv, const int* tris, int nt, float th,
                           const pbd_hash_desc& d) {
    uint64_t h = 1469598103934665603ull;
    auto mix = [&](const void* src, size_t n) {
        const unsigned char* b = (const unsigned char*)src;
        size_t i = 0;
        for (; i + 8 <= n; i += 8) {
            uint64_t w;
            memcpy(&w, b + i, 8);
            h = (h ^ w) * 1099511628211ull;
        }
        uint64_t tail = 0;
        memcpy(&tail, b + i, n - i);
        h = (h ^ tail ^ ((uint64_t)(n - i) << 56)) * 1099511628211ull;
    };
    mix(&nv, sizeof nv);
    mix(&nt, sizeof nt);
    mix(&th, sizeof th);
    mix(&d.mode, sizeof d.mode);
    mix(d.centre, sizeof d.centre);
    mix(d.size, sizeof d.size);
    mix(d.cell, sizeof d.cell);
    mix(&d.table_bits, sizeof d.table_bits);
    if (nv > 0) {
        mix(x, sizeof(float) * 3 * (size_t)nv);
        mix(p, sizeof(float) * 3 * (size_t)nv);
    }
    if (nt > 0) mix(tris, sizeof(int) * 3 * (size_t)nt);
    return h;
}

void upload_scene(pbd_ctx* c, const float* x, const float* p, const float* w, int nv, const int* tris, int nt,
                  float th = 0.0f, const pbd_hash_desc* d = nullptr) {
    PBD_CHECK(nv >= 0 && nt >= 0 && (x || !nv) && (p || !nv) && (tris || !nt), PBD_ERR_ARG, "bad scene arrays");
    for (long long i = 0; i < 3ll * nt; i++) PBD_CHECK(tris[i] >= 0 && tris[i] < nv, PBD_ERR_ARG, "triangle index out of range");
    std::vector<float4> X(nv), P(nv);
    for (int i = 0; i < nv; i++) {
        float wi = w ? w[i] : 1.0f;
        X[i] = make_float4(x[3 * i], x[3 * i + 1], x[3 * i + 2], wi);
        P[i] = make_float4(p[3 * i], p[3 * i + 1], p[3 * i + 2], wi);
    }
    c->sx.upload(X, c->stream);
    c->sp4.upload(P, c->stream);
    c->stris.upload(std::vector<int>(tris, tris + 3ll * nt), c->stream);
    c->snv = nv;
    c->snt = nt;
    c->scene_hash = scene_fingerprint(x, p, nv, tris, nt, th, d ? *d : pbd_hash_desc{});
}

// Contact detection for the whole context (all cloths, global vertex ids).
void detect_contacts(pbd_ctx* c, const float4* p) {
    hipStream_t s = c->stream;
    if (c->coll.broadphase == 1) {
        c->cbp.pairs_swept(s, c->ccfg, c->x.ptr, p, c->gtris.ptr, c->gnt, c->nv_total, c->coll.thickness);
    } else {
        c->cbp.build(s, c->ccfg, p, c->gtris.ptr, c->gnt);
        c->cbp.pairs(s, p, c->gtris.ptr, c->nv_total);
    }
    c->cbp.narrow(s, c->x.ptr, p, c->gtris.ptr, c->coll.thickness);
    if (c->coll.edge_edge) {
        // edge-edge candidates come from the centroid-cell table
        if (c->coll.broadphase == 1) c->cbp.build(s, c->ccfg, p, c->gtris.ptr, c->gnt);
        c->cbp.ee_pairs(s, p, c->gtris.ptr);
        c->cbp.ee_narrow(s, c->x.ptr, p, c->coll.thickness);
    } else {
        c->cbp.ee_npairs = 0;
        c->cbp.ee_ncontacts = 0;
    }
}

// ---- learned warm start ------------------------------------------------------
// First 6 distance prims of each vertex in prim order (the feature one-ring).
void ws_build_nbr(pbd_ctx* c) {
    if (c->ws.nbr_built) return;
    std::vector<int> nbr((size_t)6 * c->nv_total, -1);
    std::vector<float> rest((size_t)6 * c->nv_total, 0.0f);
    std::vector<unsigned char> cnt(c->nv_total, 0);
    for (auto& k : c->cloths)
        for (int p = 0; p < k.np(); p++) {
            if (k.prim_list[2 * p + 1] != 2) continue;
            int a = k.indices[k.prim_list[2 * p]] + k.vbase, b = k.indices[k.prim_list[2 * p] + 1] + k.vbase;
            if (cnt[a] < 6) { nbr[6 * (size_t)a + cnt[a]] = b; rest[6 * (size_t)a + cnt[a]] = k.rest[p]; cnt[a]++; }
            if (cnt[b] < 6) { nbr[6 * (size_t)b + cnt[b]] = a; rest[6 * (size_t)b + cnt[b]] = k.rest[p]; cnt[b]++; }
        }
    c->ws.nbr.upload(nbr, c->stream);
    c->ws.nbr_rest.upload(rest, c->stream);
    c->ws.delta.alloc(c->nv_total);
    c->ws.psave.alloc(c->nv_total);
    c->ws.rs.alloc(4 * c->cloths.size());
    c->ws.keep.alloc(c->cloths.size());
    c->ws.cap.alloc(c->cloths.size());
    c->ws.feat.alloc(ws_feature_bytes(c->nv_total) / sizeof(uint4));
    c->ws.hs.alloc(c->nv_total);
    c->ws.nbr_built = true;
    drop_graphs(c);
    if (c->ws.norm_h.empty()) {
        c->ws.norm_h.assign(47, 1.0f);
        for (int k = 0; k < 22; k++) c->ws.norm_h[k] = 0.0f;
        c->ws.norm.upload(c->ws.norm_h, c->stream);
    }
}

void ws_require_params(pbd_ctx* c) {
    PBD_CHECK(c->ws.params_ready, PBD_ERR_STATE, "warm-start weights not initialised (pbd_ws_init or pbd_ws_set_params)");
}

// Whether a warm substep of `iterations` sweeps under per-cloth budgets folds
// the keep decision into the queued tile launch's schedule (true), or makes
// it before the projection (false).
bool has_tolerance(const pbd_ctx* c);
bool queue_folds(pbd_ctx* c, int iterations) {
    if (c->sp.path != PBD_PATH_AUTO || has_tolerance(c) || c->sp.per_chunk_launches == 1) return false;
    replan_if_needed(c);
    const int ts = effective_tile_sweeps(c);
    return c->any_partial && (iterations + ts - 1) / ts > 1;
}

// Whether that launch also computes the residuals and decides
// (per_chunk_launches 2), when every multi-tile cloth's tiles fit the GPU's
// resident workgroups at once (they wait for each other): fewer tiles than
// resident slots, the grid's rounding up to a multiple of 8 not counted.
bool queue_decides(pbd_ctx* c, int iterations) {
    if (!queue_folds(c, iterations) || c->sp.per_chunk_launches != 2) return false;
    (void)walk_grid(c);  // sets walk_slots for the current launch shape
    return c->q_tiles_max < c->walk_slots;
}

// Predicts the correction and decides, per cloth, whether to keep it (its
// residual must not rise above the unwarmed start).  With apply, adds it to
// the current p of the kept cloths; otherwise the caller passes ws.delta and
// ws.keep to the tile kernel, which adds it while loading.  With advect, the
// inference kernel also advects (writes p from x and v); otherwise p holds
// the advected positions already.
// The prediction and its residuals.  defer: the keep decision is left to the
// projection (dec, filled here), which folds it into its first kernel when
// it can; otherwise it is made here (and applied to p if `apply`).
void ws_apply_prediction(pbd_ctx* c, bool apply, bool advect, float dt, QueueDecide* dec = nullptr) {
    hipStream_t s = c->stream;
    int ncl = (int)c->cloths.size();
    ws_build_nbr(c);
    float4* p = c->pbuf[c->pcur].ptr;
    // the queued tile launch may compute the residuals and decide itself;
    // then the block it needs zeroed takes the place of the residual sums
    const bool in_launch = dec && !apply && queue_decides(c, c->sp.iterations);
    unsigned long long* zero = in_launch ? c->q_zero.ptr : c->ws.rs.ptr;
    const size_t nzero = in_launch ? c->q_zero.n : 4 * (size_t)ncl;
    // the fused inference clears them itself
    const bool fused = advect;
    ws_infer(s, c->ws.W.ptr, c->cl.ptr, c->vcloth.ptr, c->x.ptr, p, advect ? c->v.ptr : nullptr, dt, c->ws.nbr.ptr,
             c->ws.nbr_rest.ptr, c->nv_total, c->ws.delta.ptr, nullptr, c->ws.norm.ptr, c->ws.feat.ptr, c->ws.hs.ptr,
             c->sp.warm_gate, fused ? zero : nullptr, fused ? (int)nzero : 0);
    if (!fused) launch_fill_u64(s, zero, nzero, 0ull);
    if (!in_launch)
        ws_resid2(s, c->rdist.ptr, (int)c->rdist.n, p, c->ws.delta.ptr, c->ws.rs.ptr, ncl, c->sp.warm_rule);
    int full = c->sp.iterations;
    int warm = c->sp.warm_iterations > 0 ? std::min(c->sp.warm_iterations, full) : full;
    if (dec && !apply) {
        *dec = QueueDecide{c->ws.rs.ptr, c->sp.warm_rule, warm, full, c->ws.keep.ptr, c->ws.cap.ptr, in_launch};
    } else {
        ws_decide(s, c->ws.rs.ptr, ncl, c->sp.warm_rule, warm, full, c->ws.cap.ptr, c->ws.keep.ptr);
        if (apply) ws_apply(s, c->vcloth.ptr, p, c->ws.delta.ptr, c->ws.keep.ptr, c->nv_total);
    }
    c->ws.used_pending = true;
}

// Per-cloth sweep counts of the last projection from the device.
void read_sweeps(pbd_ctx* c) {
    if (!c->sweeps_pending) return;
    std::vector<int> sv;
    c->sweeps.download(sv, c->stream);
    PBD_HIP(hipStreamSynchronize(c->stream));
    for (size_t ci = 0; ci < c->cloths.size(); ci++) c->cloths[ci].last_sweeps = sv[ci];
    c->sweeps_pending = false;
}

// Per-cloth record of whether the last prediction was kept, read back on
// demand so that the substep itself never waits for the host.
void ws_read_used(pbd_ctx* c) {
    if (!c->ws.used_pending) return;
    c->ws.keep.download(c->ws.used, c->stream);
    PBD_HIP(hipStreamSynchronize(c->stream));
    c->ws.used_pending = false;
}

// ---- solver orchestration ---------------------------------------------------
// True when some cloth stops on a residual tolerance (the host then reads
// residuals back between launches).
bool has_tolerance(const pbd_ctx* c) {
    if (c->sp.tol >= 0.0) return true;
    for (auto& k : c->cloths)
        if (k.tol >= 0.0) return true;
    return false;
}

// The kernel constants hold g*dt for sp.dt; recompute them when dt differs.
void advect_prepare(pbd_ctx* c, float dt) {
    if (dt != c->sp.dt) {
        c->sp.dt = dt;
        fill_cloth_dev(c);
    }
}

void advect_all(pbd_ctx* c, float dt) {
    advect_prepare(c, dt);
    launch_advect(c->stream, c->cl.ptr, c->vcloth.ptr, c->x.ptr, c->v.ptr, c->pbuf[c->pcur].ptr, 0,
                  c->nv_total, dt, 0);
    c->pfinal_valid = false;  // p is rewritten whole
}

// Bytes the tile launches of the last projection moved as designed (the
// tile_bytes profile field), from each cloth's sweep count and plan sums:
// ts sweeps per launch for multi-tile cloths (one launch for a single-tile
// one); fused_whole: the first load advects x and v; kept: the cloth's
// prediction was added on the first load; int_store: the launch holding a
// cloth's last sweep integrated it.
double tile_projection_bytes(const pbd_ctx* c, int ts, bool fused_whole, bool int_store, bool with_delta,
                             bool decide) {
    double total = 0.0;
    const int warm = c->sp.warm_iterations > 0 ? c->sp.warm_iterations : c->sp.iterations;
    for (size_t ci = 0; ci < c->cloths.size(); ci++) {
        const Cloth& k = c->cloths[ci];
        const int L = k.last_sweeps;
        if (L <= 0 || k.tile_record_bytes.size() < 2) continue;
        const bool kept = with_delta && ci < c->ws.used.size() && c->ws.used[ci] != 0;
        // owned prims are read for the residual under a tolerance, and by a
        // first chunk that decides
        const double own = (k.tol >= 0.0 ? k.tol : c->sp.tol) >= 0.0 ? 8.0 * (double)k.tiles_owned : 0.0;
        auto loaded_at = [&](int n) {
            const int row = std::min(n, (int)k.tile_record_bytes.size() - 1);
            return k.whole ? (double)k.tiles_loaded : k.tile_loaded_by_sweeps[row];
        };
        auto launch = [&](int n, bool first, bool last) {
            double load = (first && fused_whole) ? 32.0 : 16.0;
            if (first && (kept || decide)) load += 16.0;
            const bool integ = k.whole ? (fused_whole || int_store) : (int_store && last);
            const double store = integ ? 64.0 : 16.0;
            const int row = std::min(n, (int)k.tile_record_bytes.size() - 1);
            const double rec = k.whole ? n * k.tile_record_bytes[1] : k.tile_record_bytes[row];
            // a launch of n sweeps loads the prefix of each tile's list its cone needs
            const double loaded = loaded_at(n);
            double b = 4.0 * (loaded + (double)k.tiles_interior) + load * loaded + rec + own +
                       store * (double)k.tiles_interior;
            if (first && decide) {
                // the owned prims' records; a discarded cloth first loaded the
                // kept budget's cone with the prediction, then loads its own
                b += own > 0.0 ? 0.0 : 8.0 * (double)k.tiles_owned;
                if (!kept) {
                    const double lk = loaded_at(std::min(n, warm));
                    b += 36.0 * lk - 16.0 * loaded;
                }
            }
            return b;
        };
        if (k.whole) {
            total += launch(L, true, true);
        } else {
            for (int sw = 0; sw < L; sw += ts) total += launch(std::min(ts, L - sw), sw == 0, sw + ts >= L);
        }
    }
    return total;
}

// Sweeps on the current p.  Whole-cloth tiles run unfused (p_in -> p_out).
// cap: optional per-cloth sweep budgets (device array) below `iterations`.
void consolidate(pbd_ctx* c);

// Incidence lists of the Jacobi path: each vertex's prims in ascending prim
// id (distance prims, then bending prims) with its role in each.
void ensure_jacobi(pbd_ctx* c) {
    if (c->jac_built) return;
    const int nv = c->nv_total;
    std::vector<int> ddeg(nv, 0), boff((size_t)nv + 1, 0);
    for (const Cloth& k : c->cloths)
        for (int p = 0; p < k.np(); p++) {
            const int st = k.prim_list[2 * p], n = k.prim_list[2 * p + 1];
            for (int j = 0; j < n; j++) {
                const int v = k.vbase + k.indices[st + j];
                if (n == 2) ddeg[v]++;
                else boff[v + 1]++;
            }
        }
    std::vector<int> doff((size_t)nv + 1, 0);
    for (int v = 0; v < nv; v++) {
        doff[v + 1] = doff[v] + std::max(0, ddeg[v] - JAC_SLOTS);
        boff[v + 1] += boff[v];
    }
    std::vector<int2> dslot((size_t)JAC_SLOTS * nv, make_int2(0, 0)), dovf(std::max(1, doff[nv]));
    std::vector<int4> bent(std::max(1, boff[nv]));
    std::vector<int2> bmeta(std::max(1, boff[nv]));
    std::vector<int> dfill(nv, 0), bfill(boff.begin(), boff.end() - 1);
    for (const Cloth& k : c->cloths)
        for (int p = 0; p < k.np(); p++) {
            const int st = k.prim_list[2 * p], n = k.prim_list[2 * p + 1];
            const int* q = &k.indices[st];
            int bits;
            memcpy(&bits, &k.rest[p], 4);
            for (int j = 0; j < n; j++) {
                const int v = k.vbase + q[j];
                if (n == 2) {
                    int2 d = make_int2((int)((unsigned)(k.vbase + q[1 - j]) | ((unsigned)j << 31)), bits);
                    int slot = dfill[v]++;
                    if (slot < JAC_SLOTS) dslot[(size_t)slot * nv + v] = d;
                    else dovf[doff[v] + slot - JAC_SLOTS] = d;
                } else {
                    int at = bfill[v]++;
                    bent[at] = make_int4(k.vbase + q[0], k.vbase + q[1], k.vbase + q[2], k.vbase + q[3]);
                    bmeta[at] = make_int2(bits, j);
                }
            }
        }
    hipStream_t s = c->stream;
    c->jddeg.upload(ddeg, s);
    c->jdslot.upload(dslot, s);
    c->jdoff.upload(doff, s);
    c->jdovf.upload(dovf, s);
    c->jboff.upload(boff, s);
    c->jbent.upload(bent, s);
    c->jbmeta.upload(bmeta, s);
    PBD_HIP(hipStreamSynchronize(s));
    c->jac = JacobiDev{nv, c->jddeg.ptr, c->jdslot.ptr, c->jdoff.ptr, c->jdovf.ptr, c->jboff.ptr, c->jbent.ptr,
                       c->jbmeta.ptr};
    c->jac_built = true;
}

// Jacobi sweeps over vertices [vb, ve) (ping-pong between the p buffers),
// stopping early on the solver tolerance like the per-colour path.
int project_jacobi(pbd_ctx* c, int iterations, const int* cap, int vb, int ve, double tol, int ci_only) {
    hipStream_t s = c->stream;
    ensure_jacobi(c);
    const float omega = c->sp.jacobi_omega > 0.0f ? c->sp.jacobi_omega : 1.0f;
    const int ce = std::max(1, c->sp.check_every);
    const int ncl = (int)c->cloths.size();
    c->pfinal_valid = false;
    // one cloth alone: both buffers carry the other cloths' p unchanged
    if (ve - vb < c->nv_total) launch_copy(s, c->pbuf[c->pcur].ptr, c->pbuf[c->pcur ^ 1].ptr, (size_t)c->nv_total);
    int sw = 0;
    while (sw < iterations) {
        const int cur = c->pcur;
        launch_jacobi(s, c->jac, c->cl.ptr, c->vcloth.ptr, c->ktab_d.ptr, c->ktab_b.ptr, c->kmax, sw, cap, omega,
                      c->pbuf[cur].ptr, c->pbuf[cur ^ 1].ptr, vb, ve);
        c->pcur = cur ^ 1;
        sw++;
        if (tol >= 0.0 && sw % ce == 0) {
            const Cloth* one = ci_only >= 0 ? &c->cloths[ci_only] : nullptr;
            int lo = one ? one->rd_lo : 0, hi = one ? one->rd_hi : (int)c->rdist.n;
            launch_fill_u64(s, c->rmax.ptr, ncl, 0ull);
            PBD_HIP(hipMemsetAsync(c->rsum.ptr, 0, sizeof(double) * ncl, s));
            launch_residual(s, c->rdist.ptr + lo, hi - lo, c->pbuf[c->pcur].ptr, c->rmax.ptr, c->rsum.ptr,
                            nullptr);
            std::vector<unsigned long long> r;
            c->rmax.download(r, s);
            PBD_HIP(hipStreamSynchronize(s));
            double m = 0.0;
            for (auto b : r) {
                double dv;
                memcpy(&dv, &b, 8);
                m = std::max(m, dv);
            }
            if (m <= tol) break;
        }
    }
    return sw;
}

// delta / keep: a warm-start prediction the first tile launch adds while
// loading (LDS-tile path only).  integrate_end: the tiles may integrate each
// cloth when they store its last sweep; returns true if they did (the caller
// then skips integrate_all, and p is left where pfinal says).
bool project(pbd_ctx* c, int iterations, bool fused_whole, float dt, const int* cap = nullptr,
             const float4* delta = nullptr, const int* keep = nullptr, bool integrate_end = false,
             const QueueDecide* dec = nullptr) {
    hipStream_t s = c->stream;
    consolidate(c);  // an earlier substep may have left p unconsolidated
    int ncl = (int)c->cloths.size();
    double tol = c->sp.tol;
    int ce = std::max(1, c->sp.check_every);
    const bool any_tol = has_tolerance(c);
    c->sweeps_pending = false;
    // A deferred keep decision is folded into the tile queue's schedule kernel
    // (with the sweep and residual-slot clears and the final-buffer record)
    // when the queued launch runs; otherwise it is made now.
    bool fold = false;
    if (dec) {
        fold = cap && queue_folds(c, iterations);
        PBD_CHECK(!dec->in_launch || (fold && queue_decides(c, iterations)), PBD_ERR_STATE,
                  "internal: the residuals were left to a launch that does not decide");
        if (!fold) ws_decide(s, dec->rs, ncl, dec->rule, dec->warm_sweeps, dec->full_sweeps, dec->cap, dec->keep);
    }
    if (!fold) launch_fill_int(s, c->sweeps.ptr, ncl, 0);
    if (c->sp.path == PBD_PATH_JACOBI) {
        PBD_CHECK(!fused_whole && !delta, PBD_ERR_STATE, "internal: the Jacobi path is never fused");
        int sw = project_jacobi(c, iterations, cap, 0, c->nv_total, tol, -1);
        for (auto& k : c->cloths) k.last_sweeps = sw;
        if (cap) {
            launch_min_sweeps(s, c->sweeps.ptr, cap, sw, ncl);
            c->sweeps_pending = true;
        }
        return false;
    }
    if (c->sp.path == PBD_PATH_GLOBAL) {
        PBD_CHECK(!fused_whole, PBD_ERR_STATE, "internal: global path is never fused");
        PBD_CHECK(!delta, PBD_ERR_STATE, "internal: the global path applies the prediction beforehand");
        c->pfinal_valid = false;
        float4* p = c->pbuf[c->pcur].ptr;
        bool coll = false;
        for (auto& k : c->cloths) coll |= (k.prm.sphere_on || k.prm.box_on);
        int sw = 0;
        while (sw < iterations) {
            for (int col = 0; col < c->ncolors_max; col++)
                launch_project_colour(s, c->gdist.ptr, c->gd_off[col], c->gd_off[col + 1], c->gbidx.ptr, c->gbrc.ptr,
                                      c->gb_off[col], c->gb_off[col + 1], c->ktab_d.ptr, c->ktab_b.ptr, c->kmax, sw,
                                      cap, p);
            if (coll) launch_colliders(s, c->cl.ptr, c->vcloth.ptr, p, 0, c->nv_total, cap, sw);
            sw++;
            if (tol >= 0.0 && sw % ce == 0) {
                launch_fill_u64(s, c->rmax.ptr, ncl, 0ull);
                PBD_HIP(hipMemsetAsync(c->rsum.ptr, 0, sizeof(double) * ncl, s));
                launch_residual(s, c->rdist.ptr, (int)c->rdist.n, p, c->rmax.ptr, c->rsum.ptr, nullptr);
                std::vector<unsigned long long> r;
                c->rmax.download(r, s);
                PBD_HIP(hipStreamSynchronize(s));
                double m = 0.0;
                for (auto b : r) {
                    double dv;
                    memcpy(&dv, &b, 8);
                    m = std::max(m, dv);
                }
                if (m <= tol) break;
            }
        }
        for (auto& k : c->cloths) k.last_sweeps = sw;
        if (cap) {
            launch_min_sweeps(s, c->sweeps.ptr, cap, sw, ncl);
            c->sweeps_pending = true;
        }
        return false;
    }
    replan_if_needed(c);
    int ts = effective_tile_sweeps(c);
    int nchunks = (iterations + ts - 1) / ts;
    // Tiles integrate their cloth on storing its last sweep when budgets are
    // fixed and no launch that can end a cloth reads x: the advect-on-load
    // launch (the first) ends none when there are at least two chunks and no
    // per-cloth budget can end a cloth inside it.
    const bool int_store =
        integrate_end && !any_tol && !(fused_whole && c->any_partial && (nchunks < 2 || cap != nullptr));
    ensure_slots(c, nchunks);
    if (!fold) launch_fill_u64(s, c->res_slots.ptr, (size_t)c->nslots * ncl, 0ull);
    size_t lds = (size_t)std::max(1, c->plan_max_load) * sizeof(float4);
    int threads = c->sp.block_threads > 0 ? c->sp.block_threads : 1024;
    int cur = c->pcur;
    const bool bend = !c->plan.bend.empty();
    auto tiles = [&](int nblocks, const int* order, const float4* pin, float4* pout, const SolveArgs& sa,
                     const int* n_items = nullptr) {
        if (c->prof_on) PBD_HIP(hipEventRecord(next_event(c), s));
        launch_tiles(s, nblocks, order, c->tiles.ptr, c->tverts.ptr, c->tdist.ptr, c->tdist_rest.ptr, c->tbend.ptr,
                     c->tbend_rest.ptr, c->town.ptr, c->town_rest.ptr, c->tcols.ptr, c->cl.ptr, c->ktab_d.ptr,
                     c->ktab_b.ptr, c->x.ptr, c->v.ptr, pin, pout, c->res_slots.ptr, ncl, c->sweeps.ptr, cap, sa,
                     threads, lds, bend, n_items);
        if (c->prof_on) PBD_HIP(hipEventRecord(next_event(c), s));
    };
    // Under per-cloth budgets and no tolerance, a later chunk needs only the
    // tiles of cloths whose budget outlasts the first chunk (with a warm
    // start: the discarded cloths).  Their list is compacted on the device
    // once, and the later chunks walk it with a grid that fills the GPU once
    // instead of one workgroup per multi-tile tile, most of which would
    // find their cloth done.
    const bool follow = cap && !any_tol && c->any_partial && nchunks > 1;
    int follow_grid = 0;
    if (follow && c->sp.per_chunk_launches == 1) {
        launch_follow_list(s, c->order_partial.ptr, c->order_partial_cloth.ptr, (int)c->order_partial.n, cap, ts,
                           c->order_follow.ptr, c->follow_len.ptr);
        follow_grid = std::min(walk_grid(c), (int)c->order_partial.n);
    }
    SolveArgs a{};
    a.dt = dt;
    a.check_every = ce;
    a.tol = tol;
    a.kmax = c->kmax;
    a.copy_spent = any_tol ? 1 : 0;
    const int start = cur;
    int launched = 0;
    // Single-tile cloths run all their sweeps in one block and write
    // pbuf[start ^ 1]; they go into the first chunk's launch when there is one
    // (both read pbuf[start] and write pbuf[start ^ 1], on disjoint vertices).
    a.whole_nsweeps = iterations;
    a.fused = fused_whole ? 1 : 0;
    a.delta = delta;
    a.keep = keep;
    // a fused substep advects every cloth in its first launch
    a.advect_load = fused_whole ? 1 : 0;
    a.integrate_end = int_store ? 1 : 0;
    if (c->any_whole && !c->any_partial) {
        a.sweep_base = 0;
        a.nsweeps = iterations;
        a.chunk = 0;
        tiles((int)c->order_whole.n, c->order_whole.ptr, c->pbuf[cur].ptr, c->pbuf[cur ^ 1].ptr, a);
    }
    if (follow && fold && dec->in_launch) {
        // every chunk in one launch that also decides: fixed first chunks,
        // later chunks appended per discarded cloth (TileQueue)
        ensure_queue(c);
        PBD_CHECK(c->q_tile.n >= (size_t)nchunks * c->order_partial.n, PBD_ERR_STATE,
                  "internal: tile queue too small (allocated before graph capture)");
        PBD_CHECK(dec->cap == cap && delta && dec->warm_sweeps >= 1 && dec->full_sweeps == iterations && !fused_whole,
                  PBD_ERR_STATE, "internal: deciding launch arguments");
        TileQueue q{};
        q.nbr_off = c->tnbr_off.ptr;
        q.nbr = c->tnbr.ptr;
        q.pbuf0 = c->pbuf[0].ptr;
        q.pbuf1 = c->pbuf[1].ptr;
        q.start = cur;
        q.ts = ts;
        q.iterations = iterations;
        q.timeouts = c->q_timeouts.ptr;
        q.launches = c->q_timeouts.ptr + 1;
        q.stat_item = c->q_stat.ptr;
        q.n_static = (int)c->tiles.n;
        q.n_multi = c->q_n_multi;
        q.dyn_tile = c->q_tile.ptr;
        q.dyn_chunk = c->q_chunk.ptr;
        q.rs = c->q_zero.ptr;
        int* zi = reinterpret_cast<int*>(c->q_zero.ptr + (4 * (size_t)ncl + 15) / 16 * 16);
        q.ctr = zi;
        q.decided = zi + QC_COUNT;
        q.posted = q.decided + (size_t)QC_DECIDED_STRIDE * ncl;
        q.finished = q.posted + ncl;
        q.rule = dec->rule;
        q.warm_sweeps = dec->warm_sweeps;
        q.full_sweeps = dec->full_sweeps;
        q.keep = dec->keep;
        q.cap = dec->cap;
        q.pfinal = c->pfinal.ptr;
        a.sweep_base = 0;
        a.nsweeps = std::min(ts, iterations);
        a.chunk = 0;
        if (c->prof_on) PBD_HIP(hipEventRecord(next_event(c), s));
        launch_tiles_decide(s, walk_grid(c), c->tiles.ptr, c->tverts.ptr, c->tdist.ptr, c->tdist_rest.ptr,
                            c->tbend.ptr, c->tbend_rest.ptr, c->town.ptr, c->town_rest.ptr, c->tcols.ptr, c->cl.ptr,
                            c->ktab_d.ptr, c->ktab_b.ptr, c->x.ptr, c->v.ptr, c->res_slots.ptr, ncl, c->sweeps.ptr,
                            cap, a, q, threads, lds, bend);
        if (c->prof_on) PBD_HIP(hipEventRecord(next_event(c), s));
        launched = nchunks;
        if (nchunks & 1) cur ^= 1;
    } else if (follow && c->sp.per_chunk_launches != 1) {
        // every chunk in one launch: workgroups walk a queue of (tile, chunk)
        // items built on the device after the keep decision
        ensure_queue(c);
        PBD_CHECK(c->q_tile.n >= c->order_first.n + (size_t)nchunks * c->order_partial.n, PBD_ERR_STATE,
                  "internal: tile queue too small (allocated before graph capture)");
        const DBuf<int>& ord = c->any_whole ? c->order_first : c->order_partial;
        if (fold) {
            PBD_CHECK(dec->cap == cap, PBD_ERR_STATE, "internal: deferred decision for another budget array");
            launch_queue_schedule_decide(s, *dec, c->cl.ptr, ncl, dec->cap, c->sweeps.ptr, c->res_slots.ptr,
                                         c->nslots * ncl, c->pfinal.ptr, cur, iterations, ord.ptr, (int)ord.n,
                                         c->tiles.ptr, (int)c->tiles.n, ts, nchunks, c->q_tmp.ptr,
                                         c->q_tmp.ptr + c->q_tmp.n / 2, c->q_tile.ptr, c->q_chunk.ptr, c->q_count.ptr,
                                         c->q_count.ptr + 1, c->tdone.ptr);
        } else {
            launch_queue_schedule(s, ord.ptr, (int)ord.n, c->tiles.ptr, (int)c->tiles.n, cap, ts, nchunks,
                                  c->q_tmp.ptr, c->q_tmp.ptr + c->q_tmp.n / 2, c->q_tile.ptr, c->q_chunk.ptr,
                                  c->q_count.ptr, c->q_count.ptr + 1, c->tdone.ptr);
        }
        TileQueue q{};
        q.item_tile = c->q_tile.ptr;
        q.item_chunk = c->q_chunk.ptr;
        q.n_items = c->q_count.ptr;
        q.head = c->q_count.ptr + 1;
        q.done = c->tdone.ptr;
        q.nbr_off = c->tnbr_off.ptr;
        q.nbr = c->tnbr.ptr;
        q.pbuf0 = c->pbuf[0].ptr;
        q.pbuf1 = c->pbuf[1].ptr;
        q.start = cur;
        q.ts = ts;
        q.iterations = iterations;
        q.timeouts = c->q_timeouts.ptr;
        a.sweep_base = 0;
        a.nsweeps = std::min(ts, iterations);
        a.chunk = 0;
        if (c->prof_on) PBD_HIP(hipEventRecord(next_event(c), s));
        launch_tiles_queue(s, walk_grid(c), c->tiles.ptr, c->tverts.ptr, c->tdist.ptr, c->tdist_rest.ptr,
                           c->tbend.ptr, c->tbend_rest.ptr, c->town.ptr, c->town_rest.ptr, c->tcols.ptr, c->cl.ptr,
                           c->ktab_d.ptr, c->ktab_b.ptr, c->x.ptr, c->v.ptr, c->res_slots.ptr, ncl, c->sweeps.ptr,
                           cap, a, q, threads, lds, bend);
        if (c->prof_on) PBD_HIP(hipEventRecord(next_event(c), s));
        launched = nchunks;
        if (nchunks & 1) cur ^= 1;
    } else if (c->any_partial) {
        int sw = 0;
        for (int ch = 0; ch < nchunks; ch++) {
            launched++;
            a.sweep_base = sw;
            a.nsweeps = std::min(ts, iterations - sw);
            a.chunk = ch;
            if (ch > 0 && follow) {
                tiles(follow_grid, c->order_follow.ptr, c->pbuf[cur].ptr, c->pbuf[cur ^ 1].ptr, a, c->follow_len.ptr);
            } else {
                const DBuf<int>& ord = (ch == 0 && c->any_whole) ? c->order_first : c->order_partial;
                tiles((int)ord.n, ord.ptr, c->pbuf[cur].ptr, c->pbuf[cur ^ 1].ptr, a);
            }
            a.delta = nullptr;  // later chunks read what the first one wrote
            a.keep = nullptr;
            a.advect_load = 0;
            cur ^= 1;
            sw += a.nsweeps;
            if (any_tol && ch + 1 < nchunks) {
                // stop once every multi-tile cloth met the tolerance
                std::vector<unsigned long long> r(ncl);
                PBD_HIP(hipMemcpyAsync(r.data(), c->res_slots.ptr + (size_t)ch * ncl, sizeof(unsigned long long) * ncl,
                                       hipMemcpyDeviceToHost, s));
                PBD_HIP(hipStreamSynchronize(s));
                bool all = true;
                for (int ci = 0; ci < ncl; ci++) {
                    if (c->cloths[ci].whole) continue;
                    double dv;
                    memcpy(&dv, &r[ci], 8);
                    double tc = c->cloths[ci].tol >= 0.0 ? c->cloths[ci].tol : tol;
                    if (tc < 0.0 || dv > tc) all = false;
                }
                if (all) break;
            }
        }
    } else {
        cur ^= 1;
    }
    if (!fold)  // the folded schedule kernel recorded them
        launch_final_buffers(s, c->cl.ptr, ncl, start, launched, ts, iterations, any_tol ? nullptr : cap,
                             c->pfinal.ptr);
    c->pfinal_valid = true;
    c->pcur = cur;
    if (any_tol || cap) {
        // per-cloth counts live on the device; read them now only if needed
        c->sweeps_pending = true;
        if (any_tol || c->prof_on) read_sweeps(c);
    } else {
        // fixed sweeps: every cloth ran all of them
        for (auto& k : c->cloths) k.last_sweeps = iterations;
    }
    if (c->prof_on) {
        collect_events(c);
        if (delta) ws_read_used(c);
        c->prof.tile_bytes +=
            tile_projection_bytes(c, ts, fused_whole, int_store, delta != nullptr, fold && dec->in_launch);
        double bytes = 0.0;
        for (auto& k : c->cloths) {
            bytes += (double)k.last_sweeps * (kDistBytes * k.ne + kBendBytes * k.nb);
            if (k.whole && fused_whole) bytes += kVertBytes * k.nv;
            else if (fused_whole) bytes += kAdvectBytes * k.nv;  // advect on the first load
            if (int_store && !(k.whole && fused_whole)) bytes += kIntegrateBytes * k.nv;
            c->prof.vertex_sweeps += (long long)k.nv * k.last_sweeps;
            c->prof.sweeps_max = std::max<long long>(c->prof.sweeps_max, k.last_sweeps);
        }
        c->prof.tile_alg_bytes += bytes;
    }
    return int_store;
}

void integrate_all(pbd_ctx* c, float dt, bool skip_whole) {
    launch_integrate(c->stream, c->cl.ptr, c->vcloth.ptr, c->x.ptr, c->v.ptr, c->pbuf[0].ptr, c->pbuf[1].ptr,
                     c->pfinal_valid ? c->pfinal.ptr : nullptr, c->pcur, 0, c->nv_total, dt, skip_whole ? 1 : 0);
    c->pfinal_valid = false;
}

// Gathers every cloth's projected p into pbuf[pcur] (callers that read p
// between projection and integration).
void consolidate(pbd_ctx* c) {
    if (!c->pfinal_valid) return;
    launch_consolidate(c->stream, c->vcloth.ptr, c->pbuf[0].ptr, c->pbuf[1].ptr, c->pfinal.ptr, c->pcur,
                       c->nv_total);
    c->pfinal_valid = false;
}

// Sweeps ii = 0 .. n-1 of the collision substep on the current p: contacts
// detected at ii % detect_every == 0 and at ii == n - 1, one Gauss-Seidel
// sweep (cap: per-cloth budgets, a spent cloth is carried forward), the
// contact response, the analytic colliders.
void ccd_sweeps(pbd_ctx* c, int n, float dt, const int* cap) {
    hipStream_t s = c->stream;
    replan_if_needed(c);
    const int ncl = (int)c->cloths.size();
    const int every = std::max(1, c->coll.detect_every);
    ensure_slots(c, 1);
    launch_fill_u64(s, c->res_slots.ptr, (size_t)c->nslots * ncl, 0ull);
    PBD_HIP(hipMemsetAsync(c->sweeps.ptr, 0, sizeof(int) * ncl, s));
    size_t lds = (size_t)std::max(1, c->plan_max_load) * sizeof(float4);
    int threads = c->sp.block_threads > 0 ? c->sp.block_threads : 1024;
    bool coll = false;
    for (auto& k : c->cloths) coll |= (k.prm.sphere_on || k.prm.box_on);
    SolveArgs a{};
    a.dt = dt;
    a.check_every = 1;
    a.tol = -1.0;
    a.kmax = c->kmax;
    a.nsweeps = 1;
    a.whole_nsweeps = 1;
    a.fused = 0;
    a.chunk = 0;
    a.no_coll = 1;
    a.copy_spent = cap ? 1 : 0;  // a cloth past its budget is carried into the other p buffer
    c->pfinal_valid = false;
    for (int ii = 0; ii < n; ii++) {
        int cur = c->pcur;
        if (ii % every == 0 || ii == n - 1) detect_contacts(c, c->pbuf[cur].ptr);
        a.sweep_base = ii;
        if (c->order_first.n) {
            if (c->prof_on) PBD_HIP(hipEventRecord(next_event(c), s));
            launch_tiles(s, (int)c->order_first.n, c->order_first.ptr, c->tiles.ptr, c->tverts.ptr, c->tdist.ptr,
                         c->tdist_rest.ptr, c->tbend.ptr, c->tbend_rest.ptr, c->town.ptr, c->town_rest.ptr,
                         c->tcols.ptr, c->cl.ptr, c->ktab_d.ptr, c->ktab_b.ptr, c->x.ptr, c->v.ptr, c->pbuf[cur].ptr,
                         c->pbuf[cur ^ 1].ptr, c->res_slots.ptr, ncl, c->sweeps.ptr, cap, a, threads, lds,
                         !c->plan.bend.empty());
            if (c->prof_on) PBD_HIP(hipEventRecord(next_event(c), s));
        }
        c->pcur = cur ^ 1;
        c->cbp.resolve(s, c->x.ptr, c->pbuf[c->pcur].ptr, c->gtris.ptr, c->nv_total, c->coll.thickness,
                       c->coll.resolve_passes, c->coll.min_normal2 > 0.0f ? c->coll.min_normal2 : 1e-6f);
        if (coll) launch_colliders(s, c->cl.ptr, c->vcloth.ptr, c->pbuf[c->pcur].ptr, 0, c->nv_total, cap, ii);
    }
}

// Substep with collisions, in the order of projectConstraintWithColliCPU
// (PBD_cuda.cu:1130-1210): per sweep ii, [detect if ii % every == 0 or last],
// one Gauss-Seidel sweep, contact response, analytic colliders.
//
// With the warm start on, the prediction is made and decided per cloth after
// advect and added to the kept cloths' p; a kept cloth then runs
// warm_iterations sweeps, a discarded one `iterations` (exactly the plain
// collision substep).  The sweep loop runs to the largest budget of a cloth
// with prims: a cloth whose budget is spent is carried forward unprojected
// (contacts are still resolved on it), and contacts are detected every
// detect_every sweeps and at the loop's last sweep.
void step_ccd(pbd_ctx* c, float dt) {
    hipStream_t s = c->stream;
    advect_all(c, dt);
    replan_if_needed(c);
    int I = c->sp.iterations, ncl = (int)c->cloths.size();
    const int* cap = nullptr;
    std::vector<int> budget(ncl, I);
    if (c->ws.enabled) {
        ws_apply_prediction(c, true, false, dt);
        ws_read_used(c);
        const int warm = c->sp.warm_iterations > 0 ? std::min(c->sp.warm_iterations, I) : I;
        int n = 0;
        for (int ci = 0; ci < ncl; ci++) {
            budget[ci] = c->ws.used[ci] ? warm : I;
            if (c->cloths[ci].np() > 0) n = std::max(n, budget[ci]);
        }
        I = n > 0 ? n : I;
        cap = c->ws.cap.ptr;
    }
    ccd_sweeps(c, I, dt, cap);
    integrate_all(c, dt, false);
    PBD_HIP(hipStreamSynchronize(s));
    for (int ci = 0; ci < ncl; ci++) c->cloths[ci].last_sweeps = std::min(budget[ci], I);
    c->sweeps_pending = false;
    if (c->prof_on) {
        collect_events(c);
        // I launches of one sweep each, loading and storing p
        std::vector<int> ran(ncl);
        for (int ci = 0; ci < ncl; ci++) {
            ran[ci] = c->cloths[ci].last_sweeps;
            c->cloths[ci].last_sweeps = 1;
        }
        c->prof.tile_bytes += (double)I * tile_projection_bytes(c, 1, false, false, false, false);
        for (int ci = 0; ci < ncl; ci++) {
            Cloth& k = c->cloths[ci];
            k.last_sweeps = ran[ci];
            c->prof.tile_alg_bytes += (double)ran[ci] * (kDistBytes * k.ne + kBendBytes * k.nb);
            c->prof.vertex_sweeps += (long long)k.nv * ran[ci];
        }
    }
}

// One substep of the whole batch.
void substep(pbd_ctx* c, float dt) {
    // x and v hold the state; p is recomputed from them (a previous substep
    // whose tiles integrated may have left it unconsolidated: it is dead now)
    c->pfinal_valid = false;
    if (c->coll.enabled) {
        PBD_CHECK(c->sp.path != PBD_PATH_JACOBI, PBD_ERR_STATE, "the CCD pipeline runs on the Gauss-Seidel paths");
        if (c->ws.enabled) ws_require_params(c);
        step_ccd(c, dt);
    } else if (c->ws.enabled) {
        PBD_CHECK(c->sp.path != PBD_PATH_JACOBI, PBD_ERR_STATE, "the warm start runs on the Gauss-Seidel paths");
        if (c->sp.path != PBD_PATH_AUTO) {
            advect_all(c, dt);
            ws_apply_prediction(c, true, false, dt);
            project(c, c->sp.iterations, false, dt, c->ws.cap.ptr);
        } else {
            // advect inside the inference kernel; the prediction is added by the
            // tile load, and each cloth is integrated by the tiles of its last sweep
            advect_prepare(c, dt);
            QueueDecide dec{};
            ws_apply_prediction(c, false, true, dt, &dec);
            if (project(c, c->sp.iterations, false, dt, c->ws.cap.ptr, c->ws.delta.ptr, c->ws.keep.ptr, true, &dec))
                return;
        }
        integrate_all(c, dt, false);
    } else if (c->sp.path != PBD_PATH_AUTO) {
        advect_all(c, dt);
        project(c, c->sp.iterations, false, dt);
        integrate_all(c, dt, false);
    } else {
        // every tile advects on its first load and integrates on storing its last sweep
        advect_prepare(c, dt);
        if (!project(c, c->sp.iterations, true, dt, nullptr, nullptr, nullptr, true)) integrate_all(c, dt, true);
    }
}

// A substep can be replayed from a graph when it makes no decision on the
// host: fixed sweeps, no collision pipeline (its pair counts size the
// buffers), no per-launch profiling events.
bool graph_eligible(const pbd_ctx* c) {
    return c->sp.graph && !c->coll.enabled && !c->prof_on && !has_tolerance(c);
}

// Captures the substep that starts from p buffer pcur on first use, then
// replays it.  The host-side effects of a substep (which p buffer is current,
// sweep counts, the pending fallback record) are applied after every replay.
void graph_substep(pbd_ctx* c, float dt) {
    hipStream_t s = c->stream;
    const int start = c->pcur;
    if (!c->graph[start]) {
        // buffers built on first use are built before capture (it allows no allocation or wait)
        if (c->sp.path == PBD_PATH_JACOBI) ensure_jacobi(c);
        if (c->sp.path != PBD_PATH_JACOBI && c->sp.path != PBD_PATH_GLOBAL) {
            replan_if_needed(c);
            (void)walk_grid(c);  // an occupancy query, made outside the capture
            ensure_queue(c);
        }
        hipGraph_t g = nullptr;
        PBD_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
        try {
            substep(c, dt);
        } catch (...) {
            hipGraph_t junk = nullptr;
            (void)hipStreamEndCapture(s, &junk);
            if (junk) (void)hipGraphDestroy(junk);
            c->pcur = start;
            throw;
        }
        PBD_HIP(hipStreamEndCapture(s, &g));
        hipGraphExec_t ex = nullptr;
        hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        PBD_HIP(e);
        c->graph[start] = ex;
        c->graph_end[start] = c->pcur;
        c->graph_pvalid[start] = c->pfinal_valid;
        c->pcur = start;
    }
    PBD_HIP(hipGraphLaunch(c->graph[start], s));
    c->pcur = c->graph_end[start];
    c->pfinal_valid = c->graph_pvalid[start];
    if (c->ws.enabled) {
        c->ws.used_pending = true;
        c->sweeps_pending = true;
    } else {
        for (auto& k : c->cloths) k.last_sweeps = c->sp.iterations;
    }
    c->prof.graph_replays++;
}

// One substep of a single cloth, every launch restricted to its vertex and
// prim ranges: the reference's solver driven one PBDObject at a time
// (SolverPBD::SetTarget, PBD_Basic.cuh:224-233; loop body
// PositionBasedDynamics.cpp:263-290).  Same arithmetic and order as the batch.
void step_cloth(pbd_ctx* c, int ci, float dt) {
    hipStream_t s = c->stream;
    consolidate(c);  // this cloth's range is rewritten below; the others must be in place
    Cloth& k = c->cloths[ci];
    const int ncl = (int)c->cloths.size();
    const int I = c->sp.iterations, ce = std::max(1, c->sp.check_every);
    const double tol = k.tol >= 0.0 ? k.tol : c->sp.tol;
    const int vb = k.vbase, ve = k.vbase + k.nv;
    if (dt != c->sp.dt) {
        c->sp.dt = dt;
        fill_cloth_dev(c);
    }
    if (c->sp.path == PBD_PATH_JACOBI) {
        launch_advect(s, c->cl.ptr, c->vcloth.ptr, c->x.ptr, c->v.ptr, c->pbuf[c->pcur].ptr, vb, ve, dt, 0);
        k.last_sweeps = project_jacobi(c, I, nullptr, vb, ve, tol, ci);
        launch_integrate(s, c->cl.ptr, c->vcloth.ptr, c->x.ptr, c->v.ptr, c->pbuf[0].ptr, c->pbuf[1].ptr, nullptr,
                         c->pcur, vb, ve, dt, 0);
        return;
    }
    if (c->sp.path == PBD_PATH_GLOBAL) {
        float4* p = c->pbuf[c->pcur].ptr;
        launch_advect(s, c->cl.ptr, c->vcloth.ptr, c->x.ptr, c->v.ptr, p, vb, ve, dt, 0);
        const bool coll = k.prm.sphere_on || k.prm.box_on;
        int sw = 0;
        while (sw < I) {
            for (int col = 0; col < k.ncolors; col++)
                launch_project_colour(s, c->gdist.ptr, k.gd_cut[2 * col], k.gd_cut[2 * col + 1], c->gbidx.ptr,
                                      c->gbrc.ptr, k.gb_cut[2 * col], k.gb_cut[2 * col + 1], c->ktab_d.ptr,
                                      c->ktab_b.ptr, c->kmax, sw, nullptr, p);
            if (coll) launch_colliders(s, c->cl.ptr, c->vcloth.ptr, p, vb, ve);
            sw++;
            if (tol >= 0.0 && sw % ce == 0) {
                launch_fill_u64(s, c->rmax.ptr + ci, 1, 0ull);
                launch_residual(s, c->rdist.ptr + k.rd_lo, k.rd_hi - k.rd_lo, p, c->rmax.ptr, c->rsum.ptr, nullptr);
                unsigned long long r = 0;
                PBD_HIP(hipMemcpyAsync(&r, c->rmax.ptr + ci, 8, hipMemcpyDeviceToHost, s));
                PBD_HIP(hipStreamSynchronize(s));
                double m;
                memcpy(&m, &r, 8);
                if (m <= tol) break;
            }
        }
        k.last_sweeps = sw;
        launch_integrate(s, c->cl.ptr, c->vcloth.ptr, c->x.ptr, c->v.ptr, c->pbuf[0].ptr, c->pbuf[1].ptr, nullptr,
                         c->pcur, vb, ve, dt, 0);
        return;
    }
    // LDS tiles of this cloth only
    const int nblocks = k.tile_end - k.tile_begin;
    const int* order = c->tile_ids.ptr + k.tile_begin;
    int ts = effective_tile_sweeps(c);
    int nchunks = (I + ts - 1) / ts;
    ensure_slots(c, nchunks);
    launch_fill_u64(s, c->res_slots.ptr, (size_t)c->nslots * ncl, 0ull);
    launch_fill_int(s, c->sweeps.ptr + ci, 1, 0);
    size_t lds = (size_t)std::max(1, c->plan_max_load) * sizeof(float4);
    int threads = c->sp.block_threads > 0 ? c->sp.block_threads : 1024;
    SolveArgs a{};
    a.dt = dt;
    a.check_every = ce;
    a.tol = c->sp.tol;
    a.kmax = c->kmax;
    auto tiles = [&](const float4* pin, float4* pout) {
        launch_tiles(s, nblocks, order, c->tiles.ptr, c->tverts.ptr, c->tdist.ptr, c->tdist_rest.ptr, c->tbend.ptr,
                     c->tbend_rest.ptr, c->town.ptr, c->town_rest.ptr, c->tcols.ptr, c->cl.ptr, c->ktab_d.ptr,
                     c->ktab_b.ptr, c->x.ptr, c->v.ptr, pin, pout, c->res_slots.ptr, ncl, c->sweeps.ptr, nullptr, a,
                     threads, lds, !c->plan.bend.empty());
    };
    int cur = c->pcur;
    if (k.whole) {
        // advect on load, all sweeps, integrate on store: one launch
        a.nsweeps = I;
        a.whole_nsweeps = I;
        a.fused = 1;
        tiles(c->pbuf[cur ^ 1].ptr, c->pbuf[cur].ptr);  // p_in is not read when fused
    } else {
        launch_advect(s, c->cl.ptr, c->vcloth.ptr, c->x.ptr, c->v.ptr, c->pbuf[cur].ptr, vb, ve, dt, 0);
        int sw = 0;
        for (int ch = 0; ch < nchunks; ch++) {
            a.sweep_base = sw;
            a.nsweeps = std::min(ts, I - sw);
            a.chunk = ch;
            tiles(c->pbuf[cur].ptr, c->pbuf[cur ^ 1].ptr);
            cur ^= 1;
            sw += a.nsweeps;
            if (tol >= 0.0 && ch + 1 < nchunks) {
                unsigned long long r = 0;
                PBD_HIP(hipMemcpyAsync(&r, c->res_slots.ptr + (size_t)ch * ncl + ci, 8, hipMemcpyDeviceToHost, s));
                PBD_HIP(hipStreamSynchronize(s));
                double m;
                memcpy(&m, &r, 8);
                if (m <= tol) break;
            }
        }
        // the batch's current p buffer must hold this cloth's result
        if (cur != c->pcur) launch_copy(s, c->pbuf[cur].ptr + vb, c->pbuf[c->pcur].ptr + vb, k.nv);
        launch_integrate(s, c->cl.ptr, c->vcloth.ptr, c->x.ptr, c->v.ptr, c->pbuf[0].ptr, c->pbuf[1].ptr, nullptr,
                         c->pcur, vb, ve, dt, 0);
    }
    if (tol >= 0.0) {
        int sv = 0;
        PBD_HIP(hipMemcpyAsync(&sv, c->sweeps.ptr + ci, sizeof(int), hipMemcpyDeviceToHost, s));
        PBD_HIP(hipStreamSynchronize(s));
        k.last_sweeps = sv;
    } else {
        k.last_sweeps = I;
    }
}

// pbd_step's loop: graph replays where eligible, eager substeps otherwise.
void run_substeps(pbd_ctx* c, int n) {
    replan_if_needed(c);
    const float dt = c->sp.dt;
    const bool graph = graph_eligible(c);
    if (graph) {
        // everything a capture must not do (allocations, uploads) happens first
        int ts = effective_tile_sweeps(c);
        ensure_slots(c, (c->sp.iterations + ts - 1) / ts);
        if (c->ws.enabled) ws_build_nbr(c);
    }
    for (int s = 0; s < n; s++) {
        if (graph)
            graph_substep(c, dt);
        else
            substep(c, dt);
    }
}

}  // namespace

// ============================================================================
extern "C" {

const char* pbd_last_error(void) { return g_err.c_str(); }

int pbd_create(int device, pbd_ctx** out) {
    return guard([&] {
        PBD_CHECK(out, PBD_ERR_ARG, "null output pointer");
        *out = nullptr;
        int n = 0;
        PBD_HIP(hipGetDeviceCount(&n));
        PBD_CHECK(device >= 0 && device < n, PBD_ERR_ARG, "device index out of range");
        auto* c = new pbd_ctx();
        c->device = device;
        try {
            PBD_HIP(hipSetDevice(device));
            PBD_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
            int lds = 0;
            PBD_HIP(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
            PBD_HIP(hipDeviceGetAttribute(&c->num_cu, hipDeviceAttributeMultiprocessorCount, device));
            // lds_max - TILE_STATIC_LDS is what the tile kernel's dynamic LDS may reach
            const size_t dyn = tiles_lds_optin((size_t)std::max(lds, 65536) - TILE_STATIC_LDS);
            c->lds_max = (int)dyn + TILE_STATIC_LDS;
        } catch (...) {
            delete c;
            throw;
        }
        c->sp.dt = 1.0f / 600.0f;
        c->sp.iterations = 20;
        c->sp.tol = -1.0;
        c->sp.check_every = 1;
        c->sp.path = PBD_PATH_AUTO;
        *out = c;
    });
}

int pbd_destroy(pbd_ctx* ctx) {
    return guard([&] {
        if (!ctx) return;
        (void)hipSetDevice(ctx->device);
        if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
        hipStream_t s = ctx->stream;
        delete ctx;
        if (s) (void)hipStreamDestroy(s);
    });
}

int pbd_synchronize(pbd_ctx* ctx) {
    return guard([&] {
        PBD_CHECK(ctx, PBD_ERR_ARG, "null context");
        PBD_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int pbd_cloth_create_grid(pbd_ctx* ctx, const pbd_grid_desc* d, int* cloth) {
    return guard([&] {
        PBD_CHECK(ctx && d && cloth, PBD_ERR_ARG, "null argument");
        PBD_CHECK(!ctx->inited, PBD_ERR_STATE, "cloths must be created before pbd_init");
        PBD_CHECK(d->res_x >= 2 && d->res_y >= 2, PBD_ERR_ARG, "grid resolution must be >= 2");
        PBD_CHECK((long long)d->res_x * d->res_y < (1ll << 30), PBD_ERR_LIMIT, "grid too large");
        Cloth c;
        c.rx = d->res_x;
        c.ry = d->res_y;
        make_grid(*d, c.pos0, c.tris);
        c.nv = c.rx * c.ry;
        c.nt = (int)c.tris.size() / 3;
        c.prm = default_params();
        ctx->cloths.push_back(std::move(c));
        *cloth = (int)ctx->cloths.size() - 1;
    });
}

int pbd_cloth_create_mesh(pbd_ctx* ctx, const float* pos, int nv, const int* tris, int nt, int* cloth) {
    return guard([&] {
        PBD_CHECK(ctx && cloth && (pos || nv == 0) && (tris || nt == 0), PBD_ERR_ARG, "null argument");
        PBD_CHECK(!ctx->inited, PBD_ERR_STATE, "cloths must be created before pbd_init");
        PBD_CHECK(nv >= 0 && nt >= 0, PBD_ERR_ARG, "negative size");
        Cloth c;
        c.pos0.assign(pos, pos + (size_t)3 * nv);
        c.tris.assign(tris, tris + (size_t)3 * nt);
        for (int i : c.tris) PBD_CHECK(i >= 0 && i < nv, PBD_ERR_ARG, "triangle index out of range");
        c.nv = nv;
        c.nt = nt;
        c.prm = default_params();
        ctx->cloths.push_back(std::move(c));
        *cloth = (int)ctx->cloths.size() - 1;
    });
}

int pbd_cloth_load_txt(pbd_ctx* ctx, const char* topol, const char* constr, int* cloth) {
    return guard([&] {
        PBD_CHECK(ctx && topol && cloth, PBD_ERR_ARG, "null argument");
        PBD_CHECK(!ctx->inited, PBD_ERR_STATE, "cloths must be created before pbd_init");
        Cloth c;
        read_txt_cloth(topol, constr, c.pos0, c.tris, c.user_pairs, c.has_user_pairs);
        c.nv = (int)c.pos0.size() / 3;
        c.nt = (int)c.tris.size() / 3;
        for (int i : c.tris) PBD_CHECK(i >= 0 && i < c.nv, PBD_ERR_IO, "triangle index out of range");
        c.prm = default_params();
        ctx->cloths.push_back(std::move(c));
        *cloth = (int)ctx->cloths.size() - 1;
    });
}

int pbd_cloth_load_cache(pbd_ctx* ctx, const char* mesh_path, const char* constr_path, int* cloth) {
    return guard([&] {
        PBD_CHECK(ctx && mesh_path && cloth, PBD_ERR_ARG, "null argument");
        PBD_CHECK(!ctx->inited, PBD_ERR_STATE, "cloths must be created before pbd_init");
        Cloth c;
        c.prm = default_params();
        read_cache(mesh_path, constr_path, c);
        ctx->cloths.push_back(std::move(c));
        *cloth = (int)ctx->cloths.size() - 1;
    });
}

int pbd_cloth_save_cache(pbd_ctx* ctx, int id, const char* mesh_path, const char* constr_path, int precision) {
    return guard([&] {
        Cloth& c = cloth_of(ctx, id);
        PBD_CHECK(ctx->inited, PBD_ERR_STATE, "call pbd_init first");
        PBD_CHECK(precision >= 0 && precision <= 17, PBD_ERR_ARG, "precision must be 0 (6 digits) to 17");
        set_device(ctx);
        consolidate(ctx);
        std::vector<float> st[3];
        const DBuf<float4>* src[3] = {&ctx->x, &ctx->v, &ctx->pbuf[ctx->pcur]};
        std::vector<float4> h(c.nv);
        for (int k = 0; k < 3; k++) {
            PBD_HIP(hipMemcpyAsync(h.data(), src[k]->ptr + c.vbase, sizeof(float4) * c.nv, hipMemcpyDeviceToHost,
                                   ctx->stream));
            PBD_HIP(hipStreamSynchronize(ctx->stream));
            st[k].resize((size_t)3 * c.nv);
            for (int i = 0; i < c.nv; i++) {
                st[k][3 * i] = h[i].x;
                st[k][3 * i + 1] = h[i].y;
                st[k][3 * i + 2] = h[i].z;
            }
        }
        write_cache(c, st[0].data(), st[1].data(), st[2].data(), mesh_path, constr_path, precision);
    });
}

int pbd_cloth_set_distance_pairs(pbd_ctx* ctx, int id, const int* pairs, int ne) {
    return guard([&] {
        Cloth& c = cloth_of(ctx, id);
        PBD_CHECK(!ctx->inited, PBD_ERR_STATE, "constraints are fixed after pbd_init");
        PBD_CHECK(ne >= 0 && (pairs || ne == 0), PBD_ERR_ARG, "bad pair array");
        c.user_pairs.assign(pairs, pairs + (size_t)2 * ne);
        c.has_user_pairs = true;
    });
}

int pbd_cloth_set_constraint_option(pbd_ctx* ctx, int id, unsigned ct, const float* st) {
    return guard([&] {
        Cloth& c = cloth_of(ctx, id);
        PBD_CHECK(!ctx->inited, PBD_ERR_STATE, "constraints are fixed after pbd_init");
        PBD_CHECK(!(ct & ~(unsigned)(PBD_DISTANCE | PBD_BENDING | PBD_ANCHOR)), PBD_ERR_ARG, "unknown constraint bits");
        c.ct = ct;
        if (st)
            for (int k = 0; k < 3; k++) {
                PBD_CHECK(st[k] >= 0.0f && st[k] <= 1.0f, PBD_ERR_ARG, "stiffness must be in [0,1]");
                c.stiff[k] = st[k];
            }
    });
}

int pbd_cloth_set_pins(pbd_ctx* ctx, int id, const int* verts, int n) {
    return guard([&] {
        Cloth& c = cloth_of(ctx, id);
        PBD_CHECK(!ctx->inited, PBD_ERR_STATE, "pins are fixed after pbd_init");
        PBD_CHECK(n >= 0 && (verts || n == 0), PBD_ERR_ARG, "bad pin array");
        for (int i = 0; i < n; i++) PBD_CHECK(verts[i] >= 0 && verts[i] < c.nv, PBD_ERR_ARG, "pin out of range");
        c.pins.assign(verts, verts + n);
    });
}

int pbd_cloth_set_params(pbd_ctx* ctx, int id, const pbd_cloth_params* p) {
    return guard([&] {
        Cloth& c = cloth_of(ctx, id);
        PBD_CHECK(p, PBD_ERR_ARG, "null params");
        PBD_CHECK(p->damping > 0.0f, PBD_ERR_ARG, "damping must be positive");
        PBD_CHECK(!p->sphere_on || p->sphere_r > 0.0f, PBD_ERR_ARG, "sphere radius must be positive");
        c.prm = *p;
        if (ctx->inited) {
            set_device(ctx);
            fill_cloth_dev(ctx);
        }
    });
}

int pbd_cloth_get_params(pbd_ctx* ctx, int id, pbd_cloth_params* p) {
    return guard([&] {
        Cloth& c = cloth_of(ctx, id);
        PBD_CHECK(p, PBD_ERR_ARG, "null params");
        *p = c.prm;
    });
}

int pbd_init(pbd_ctx* ctx, int color_rounds) {
    return guard([&] {
        PBD_CHECK(ctx, PBD_ERR_ARG, "null context");
        PBD_CHECK(!ctx->inited, PBD_ERR_STATE, "pbd_init already called");
        do_init(ctx, color_rounds);
    });
}

int pbd_cloth_get_info(pbd_ctx* ctx, int id, pbd_cloth_info* info) {
    return guard([&] {
        Cloth& c = cloth_of(ctx, id);
        PBD_CHECK(info, PBD_ERR_ARG, "null info");
        PBD_CHECK(ctx->inited, PBD_ERR_STATE, "call pbd_init first");
        info->nv = c.nv;
        info->nt = c.nt;
        info->ne = c.ne;
        info->nb = c.nb;
        info->nprims = c.np();
        info->ncolors = c.ncolors;
        info->color_rounds = c.color_rounds;
        info->adj_len = (int)c.adj.size();
        info->ntiles = c.tile_end - c.tile_begin;
        info->whole = c.whole ? 1 : 0;
        info->max_loaded = c.max_loaded;
    });
}

int pbd_cloth_get_prims(pbd_ctx* ctx, int id, int* prim_list, int* indices, float* rest) {
    return guard([&] {
        Cloth& c = cloth_of(ctx, id);
        PBD_CHECK(ctx->inited, PBD_ERR_STATE, "call pbd_init first");
        if (prim_list) std::copy(c.prim_list.begin(), c.prim_list.end(), prim_list);
        if (indices) std::copy(c.indices.begin(), c.indices.end(), indices);
        if (rest) std::copy(c.rest.begin(), c.rest.end(), rest);
    });
}

int pbd_cloth_get_adjacency(pbd_ctx* ctx, int id, int* start_num, int* adj) {
    return guard([&] {
        Cloth& c = cloth_of(ctx, id);
        PBD_CHECK(ctx->inited, PBD_ERR_STATE, "call pbd_init first");
        if (start_num) std::copy(c.adj_start_num.begin(), c.adj_start_num.end(), start_num);
        if (adj) std::copy(c.adj.begin(), c.adj.end(), adj);
    });
}

int pbd_cloth_get_colors(pbd_ctx* ctx, int id, int* color, int* sorted, int* worksets) {
    return guard([&] {
        Cloth& c = cloth_of(ctx, id);
        PBD_CHECK(ctx->inited, PBD_ERR_STATE, "call pbd_init first");
        if (color) std::copy(c.color.begin(), c.color.end(), color);
        if (sorted) std::copy(c.sorted.begin(), c.sorted.end(), sorted);
        if (worksets) std::copy(c.worksets.begin(), c.worksets.end(), worksets);
    });
}

int pbd_set_solver_params(pbd_ctx* ctx, const pbd_solver_params* p) {
    return guard([&] {
        PBD_CHECK(ctx && p, PBD_ERR_ARG, "null argument");
        PBD_CHECK(p->dt > 0.0f, PBD_ERR_ARG, "dt must be positive");
        PBD_CHECK(p->iterations >= 1 && p->iterations <= 1000000, PBD_ERR_ARG, "iterations out of range");
        PBD_CHECK(p->check_every >= 1, PBD_ERR_ARG, "check_every must be >= 1");
        PBD_CHECK(p->path == PBD_PATH_AUTO || p->path == PBD_PATH_GLOBAL || p->path == PBD_PATH_JACOBI, PBD_ERR_ARG,
                  "unknown path");
        PBD_CHECK(p->jacobi_omega >= 0.0f && p->jacobi_omega < 2.0f, PBD_ERR_ARG, "jacobi_omega must be in [0, 2)");
        PBD_CHECK(p->split_whole == 0 || p->split_whole == 1, PBD_ERR_ARG, "split_whole must be 0 or 1");
        PBD_CHECK(p->per_chunk_launches >= 0 && p->per_chunk_launches <= 2, PBD_ERR_ARG,
                  "per_chunk_launches must be 0, 1 or 2");
        PBD_CHECK(p->tile_sweeps >= 0 && p->tile_max_verts >= 0, PBD_ERR_ARG, "negative tile setting");
        PBD_CHECK(p->block_threads == 0 || (p->block_threads % 64 == 0 && p->block_threads <= 1024), PBD_ERR_ARG,
                  "block_threads must be a multiple of 64 up to 1024");
        PBD_CHECK(p->warm_iterations >= 0, PBD_ERR_ARG, "warm_iterations must be >= 0");
        PBD_CHECK(p->warm_rule == 0 || p->warm_rule == 1, PBD_ERR_ARG, "warm_rule must be 0 (max) or 1 (RMS)");
        PBD_CHECK(p->warm_gate >= 0.0f, PBD_ERR_ARG, "warm_gate must be >= 0");
        ctx->sp = *p;
        if (ctx->inited) {
            set_device(ctx);
            fill_ktabs(ctx);
            fill_cloth_dev(ctx);
            replan_if_needed(ctx);
        }
    });
}

int pbd_advect(pbd_ctx* ctx, float dt) {
    return guard([&] {
        PBD_CHECK(ctx && ctx->inited, PBD_ERR_STATE, "call pbd_init first");
        PBD_CHECK(dt > 0.0f, PBD_ERR_ARG, "dt must be positive");
        set_device(ctx);
        advect_all(ctx, dt);
    });
}

int pbd_project(pbd_ctx* ctx, int iterations) {
    return guard([&] {
        PBD_CHECK(ctx && ctx->inited, PBD_ERR_STATE, "call pbd_init first");
        PBD_CHECK(iterations >= 1, PBD_ERR_ARG, "iterations must be >= 1");
        set_device(ctx);
        if (iterations > ctx->kmax) {
            ctx->sp.iterations = iterations;
            fill_ktabs(ctx);
            fill_cloth_dev(ctx);
        }
        project(ctx, iterations, false, ctx->sp.dt);
        consolidate(ctx);
    });
}

int pbd_integrate(pbd_ctx* ctx, float dt) {
    return guard([&] {
        PBD_CHECK(ctx && ctx->inited, PBD_ERR_STATE, "call pbd_init first");
        PBD_CHECK(dt > 0.0f, PBD_ERR_ARG, "dt must be positive");
        set_device(ctx);
        integrate_all(ctx, dt, false);
    });
}

int pbd_step(pbd_ctx* ctx, int substeps) {
    return guard([&] {
        PBD_CHECK(ctx && ctx->inited, PBD_ERR_STATE, "call pbd_init first");
        PBD_CHECK(substeps >= 0, PBD_ERR_ARG, "substeps must be >= 0");
        set_device(ctx);
        auto t0 = std::chrono::steady_clock::now();
        run_substeps(ctx, substeps);
        if (ctx->prof_on) {
            PBD_HIP(hipStreamSynchronize(ctx->stream));
            ctx->prof.step_ms +=
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            ctx->prof.substeps += substeps;
        }
    });
}

int pbd_step_cloth(pbd_ctx* ctx, int cloth, int substeps) {
    return guard([&] {
        PBD_CHECK(ctx && ctx->inited, PBD_ERR_STATE, "call pbd_init first");
        cloth_of(ctx, cloth);
        PBD_CHECK(substeps >= 0, PBD_ERR_ARG, "substeps must be >= 0");
        PBD_CHECK(!ctx->coll.enabled, PBD_ERR_STATE, "collisions are solved for the whole batch (use pbd_step)");
        PBD_CHECK(!ctx->ws.enabled, PBD_ERR_STATE, "the warm start runs on the whole batch (use pbd_step)");
        set_device(ctx);
        replan_if_needed(ctx);
        for (int s = 0; s < substeps; s++) step_cloth(ctx, cloth, ctx->sp.dt);
    });
}

int pbd_sh_build(pbd_ctx* ctx, const pbd_hash_desc* d, const float* pos, int nv, const int* tris, int nt) {
    return guard([&] {
        PBD_CHECK(ctx && d, PBD_ERR_ARG, "null argument");
        set_device(ctx);
        HashCfg g = make_hash_cfg(*d);
        upload_scene(ctx, pos, pos, nullptr, nv, tris, nt);
        ctx->sbp.build(ctx->stream, g, ctx->sp4.ptr, ctx->stris.ptr, nt);
        PBD_HIP(hipStreamSynchronize(ctx->stream));
        ctx->sh_built = true;
        ctx->ccd_built = false;
    });
}

int pbd_sh_build_noalloc(int device, void* stream, const pbd_hash_desc* d, const float* pos, int nv, const int* tris,
                         int nt, int* sorted_keys, int* sorted_ids, int* raw_keys, int* cell_start, int* cell_end) {
    return guard([&] {
        PBD_CHECK(d, PBD_ERR_ARG, "null grid description");
        PBD_CHECK(nv >= 0 && nt >= 0, PBD_ERR_ARG, "negative vertex or triangle count");
        PBD_CHECK(nt <= SH_NOALLOC_MAX_TRIS, PBD_ERR_LIMIT,
                  "pbd_sh_build_noalloc: " + std::to_string(nt) + " triangles exceed its maximum of " +
                      std::to_string(SH_NOALLOC_MAX_TRIS));
        PBD_CHECK(nt == 0 || (pos && tris && sorted_keys && sorted_ids && raw_keys), PBD_ERR_ARG,
                  "null device array");
        PBD_CHECK(cell_start && cell_end, PBD_ERR_ARG, "null cell range array");
        const HashCfg g = make_hash_cfg(*d);
        int n = 0;
        PBD_HIP(hipGetDeviceCount(&n));
        PBD_CHECK(device >= 0 && device < n, PBD_ERR_ARG, "device index out of range");
        // work on `device`, then give the calling thread its current device back
        int prev = 0;
        PBD_HIP(hipGetDevice(&prev));
        PBD_HIP(hipSetDevice(device));
        struct Back {
            int dev;
            ~Back() { (void)hipSetDevice(dev); }
        } back{prev};
        sh_build_noalloc((hipStream_t)stream, device, g, pos, nv, tris, nt, sorted_keys, sorted_ids, raw_keys,
                         cell_start, cell_end);
    });
}

int pbd_sh_num_cells(pbd_ctx* ctx, int* ncells) {
    return guard([&] {
        PBD_CHECK(ctx && ncells, PBD_ERR_ARG, "null argument");
        PBD_CHECK(ctx->sh_built, PBD_ERR_STATE, "call pbd_sh_build first");
        *ncells = ctx->sbp.cfg.ncells;
    });
}

int pbd_sh_get(pbd_ctx* ctx, int* keys, int* ids, int* raw, int* cstart, int* cend) {
    return guard([&] {
        PBD_CHECK(ctx, PBD_ERR_ARG, "null context");
        PBD_CHECK(ctx->sh_built, PBD_ERR_STATE, "call pbd_sh_build first");
        set_device(ctx);
        hipStream_t s = ctx->stream;
        Broadphase& b = ctx->sbp;
        int nt = b.nt;
        if (nt > 0) {
            std::vector<uint32_t> k1;
            if (keys) {
                b.keys.download(k1, s);
                PBD_HIP(hipStreamSynchronize(s));
                for (int i = 0; i < nt; i++) keys[i] = (int)k1[i] - 1;
            }
            if (ids) PBD_HIP(hipMemcpyAsync(ids, b.ids.ptr, sizeof(int) * nt, hipMemcpyDeviceToHost, s));
            if (raw) PBD_HIP(hipMemcpyAsync(raw, b.raw.ptr, sizeof(int) * nt, hipMemcpyDeviceToHost, s));
        }
        if (cstart) PBD_HIP(hipMemcpyAsync(cstart, b.cstart.ptr, sizeof(int) * b.cfg.ncells, hipMemcpyDeviceToHost, s));
        if (cend) PBD_HIP(hipMemcpyAsync(cend, b.cend.ptr, sizeof(int) * b.cfg.ncells, hipMemcpyDeviceToHost, s));
        PBD_HIP(hipStreamSynchronize(s));
    });
}

int pbd_sh_pairs(pbd_ctx* ctx, long long cap, int64_t* out, long long* count) {
    return guard([&] {
        PBD_CHECK(ctx && count, PBD_ERR_ARG, "null argument");
        PBD_CHECK(ctx->sh_built, PBD_ERR_STATE, "call pbd_sh_build first");
        set_device(ctx);
        hipStream_t s = ctx->stream;
        long long n = ctx->sbp.pairs(s, ctx->sp4.ptr, ctx->stris.ptr, ctx->snv);
        *count = n;
        if (out && n > 0) {
            long long m = std::min(n, cap);
            std::vector<uint32_t> v(m), f(m);
            PBD_HIP(hipMemcpyAsync(v.data(), ctx->sbp.uv.ptr, 4 * m, hipMemcpyDeviceToHost, s));
            PBD_HIP(hipMemcpyAsync(f.data(), ctx->sbp.uf.ptr, 4 * m, hipMemcpyDeviceToHost, s));
            PBD_HIP(hipStreamSynchronize(s));
            for (long long i = 0; i < m; i++) out[i] = ((int64_t)v[i] << 32) | (int64_t)f[i];
        }
    });
}

int pbd_ccd_detect(pbd_ctx* ctx, const pbd_hash_desc* d, const float* x, const float* p, int nv, const int* tris,
                   int nt, float th, long long* nc) {
    return guard([&] {
        PBD_CHECK(ctx && d && nc, PBD_ERR_ARG, "null argument");
        PBD_CHECK(th >= 0.0f, PBD_ERR_ARG, "thickness must be >= 0");
        set_device(ctx);
        hipStream_t s = ctx->stream;
        HashCfg g = make_hash_cfg(*d);
        upload_scene(ctx, x, p, nullptr, nv, tris, nt, th, d);
        ctx->sbp.build(s, g, ctx->sp4.ptr, ctx->stris.ptr, nt);
        ctx->sbp.pairs(s, ctx->sp4.ptr, ctx->stris.ptr, nv);
        *nc = ctx->sbp.narrow(s, ctx->sx.ptr, ctx->sp4.ptr, ctx->stris.ptr, th);
        ctx->sbp.ee_npairs = 0;
        ctx->sbp.ee_ncontacts = 0;
        ctx->sh_built = true;
        ctx->ccd_built = true;
    });
}

int pbd_ccd_detect_ee(pbd_ctx* ctx, const pbd_hash_desc* d, const float* x, const float* p, int nv, const int* tris,
                      int nt, float th, long long* npairs, long long* nc) {
    return guard([&] {
        PBD_CHECK(ctx && d && nc, PBD_ERR_ARG, "null argument");
        PBD_CHECK(th >= 0.0f, PBD_ERR_ARG, "thickness must be >= 0");
        set_device(ctx);
        hipStream_t s = ctx->stream;
        HashCfg g = make_hash_cfg(*d);
        const uint64_t prev_hash = ctx->scene_hash;
        // vertex-face contacts of an earlier detection on another scene, or
        // at another thickness or grid, do not carry over
        upload_scene(ctx, x, p, nullptr, nv, tris, nt, th, d);
        bool same = ctx->ccd_built && ctx->scene_hash == prev_hash;
        if (!same) {
            ctx->sbp.npairs = 0;
            ctx->sbp.ncontacts = 0;
        }
        std::vector<int> te;
        std::vector<int2> ed;
        build_edge_table(tris, nt, te, ed);
        ctx->sbp.set_edges(s, te, ed);
        ctx->sbp.build(s, g, ctx->sp4.ptr, ctx->stris.ptr, nt);
        long long np = ctx->sbp.ee_pairs(s, ctx->sp4.ptr, ctx->stris.ptr);
        if (npairs) *npairs = np;
        *nc = ctx->sbp.ee_narrow(s, ctx->sx.ptr, ctx->sp4.ptr, th);
        ctx->sh_built = true;
        ctx->ccd_built = true;
    });
}

int pbd_ccd_get_ee_pairs(pbd_ctx* ctx, long long cap, int* quads, long long* count) {
    return guard([&] {
        PBD_CHECK(ctx && count, PBD_ERR_ARG, "null argument");
        PBD_CHECK(ctx->ccd_built, PBD_ERR_STATE, "call pbd_ccd_detect_ee first");
        set_device(ctx);
        hipStream_t s = ctx->stream;
        Broadphase& b = ctx->sbp;
        *count = b.ee_npairs;
        long long m = std::min(cap, b.ee_npairs);
        if (m <= 0 || !quads) return;
        std::vector<uint32_t> ea(m), eb(m);
        std::vector<int2> ed;
        PBD_HIP(hipMemcpyAsync(ea.data(), b.eua.ptr, 4 * m, hipMemcpyDeviceToHost, s));
        PBD_HIP(hipMemcpyAsync(eb.data(), b.eub.ptr, 4 * m, hipMemcpyDeviceToHost, s));
        b.edges.download(ed, s);
        PBD_HIP(hipStreamSynchronize(s));
        for (long long i = 0; i < m; i++) {
            int2 a = ed[ea[i]], c = ed[eb[i]];
            quads[4 * i] = a.x;
            quads[4 * i + 1] = a.y;
            quads[4 * i + 2] = c.x;
            quads[4 * i + 3] = c.y;
        }
    });
}

int pbd_ccd_get_ee_contacts(pbd_ctx* ctx, long long cap, int* quads, float* t, float* n, double* w,
                            long long* count) {
    return guard([&] {
        PBD_CHECK(ctx, PBD_ERR_ARG, "null context");
        PBD_CHECK(ctx->ccd_built, PBD_ERR_STATE, "call pbd_ccd_detect_ee first");
        set_device(ctx);
        hipStream_t s = ctx->stream;
        Broadphase& b = ctx->sbp;
        if (count) *count = b.ee_ncontacts;
        long long m = std::min(cap, b.ee_ncontacts);
        if (m <= 0) return;
        if (quads) PBD_HIP(hipMemcpyAsync(quads, b.ecq.ptr, 16 * m, hipMemcpyDeviceToHost, s));
        if (t) PBD_HIP(hipMemcpyAsync(t, b.ect.ptr, 4 * m, hipMemcpyDeviceToHost, s));
        if (n) PBD_HIP(hipMemcpyAsync(n, b.ecn.ptr, 12 * m, hipMemcpyDeviceToHost, s));
        if (w) PBD_HIP(hipMemcpyAsync(w, b.ecw.ptr, 32 * m, hipMemcpyDeviceToHost, s));
        PBD_HIP(hipStreamSynchronize(s));
    });
}

int pbd_ccd_detect_swept(pbd_ctx* ctx, const pbd_hash_desc* d, const float* x, const float* p, int nv,
                         const int* tris, int nt, float th, long long* npairs, long long* nc) {
    return guard([&] {
        PBD_CHECK(ctx && d && nc, PBD_ERR_ARG, "null argument");
        PBD_CHECK(th >= 0.0f, PBD_ERR_ARG, "thickness must be >= 0");
        set_device(ctx);
        hipStream_t s = ctx->stream;
        HashCfg g = make_hash_cfg(*d);
        upload_scene(ctx, x, p, nullptr, nv, tris, nt, th, d);
        long long np = ctx->sbp.pairs_swept(s, g, ctx->sx.ptr, ctx->sp4.ptr, ctx->stris.ptr, nt, nv, th);
        if (npairs) *npairs = np;
        *nc = ctx->sbp.narrow(s, ctx->sx.ptr, ctx->sp4.ptr, ctx->stris.ptr, th);
        ctx->sbp.ee_npairs = 0;
        ctx->sbp.ee_ncontacts = 0;
        ctx->sh_built = false;  // the cell table holds swept entries, not centroid keys
        ctx->ccd_built = true;
    });
}

int pbd_ccd_get_pairs(pbd_ctx* ctx, long long cap, int64_t* out, long long* count) {
    return guard([&] {
        PBD_CHECK(ctx && count, PBD_ERR_ARG, "null argument");
        PBD_CHECK(ctx->ccd_built, PBD_ERR_STATE, "call pbd_ccd_detect or pbd_ccd_detect_swept first");
        set_device(ctx);
        hipStream_t s = ctx->stream;
        const long long n = ctx->sbp.npairs;
        *count = n;
        if (out && n > 0) {
            long long m = std::min(n, cap);
            std::vector<uint32_t> v(m), f(m);
            PBD_HIP(hipMemcpyAsync(v.data(), ctx->sbp.uv.ptr, 4 * m, hipMemcpyDeviceToHost, s));
            PBD_HIP(hipMemcpyAsync(f.data(), ctx->sbp.uf.ptr, 4 * m, hipMemcpyDeviceToHost, s));
            PBD_HIP(hipStreamSynchronize(s));
            for (long long i = 0; i < m; i++) out[i] = ((int64_t)v[i] << 32) | (int64_t)f[i];
        }
    });
}

int pbd_ccd_get_contacts(pbd_ctx* ctx, long long cap, int* vf, float* t, float* n, double* w) {
    return guard([&] {
        PBD_CHECK(ctx, PBD_ERR_ARG, "null context");
        PBD_CHECK(ctx->ccd_built, PBD_ERR_STATE, "call pbd_ccd_detect first");
        set_device(ctx);
        hipStream_t s = ctx->stream;
        Broadphase& b = ctx->sbp;
        long long m = std::min(cap, b.ncontacts);
        if (m <= 0) return;
        if (vf) {
            std::vector<int> v(m), f(m);
            PBD_HIP(hipMemcpyAsync(v.data(), b.cv.ptr, 4 * m, hipMemcpyDeviceToHost, s));
            PBD_HIP(hipMemcpyAsync(f.data(), b.cf.ptr, 4 * m, hipMemcpyDeviceToHost, s));
            PBD_HIP(hipStreamSynchronize(s));
            for (long long i = 0; i < m; i++) {
                vf[2 * i] = v[i];
                vf[2 * i + 1] = f[i];
            }
        }
        if (t) PBD_HIP(hipMemcpyAsync(t, b.ct.ptr, 4 * m, hipMemcpyDeviceToHost, s));
        if (n) PBD_HIP(hipMemcpyAsync(n, b.cn.ptr, 12 * m, hipMemcpyDeviceToHost, s));
        if (w) PBD_HIP(hipMemcpyAsync(w, b.cw.ptr, 32 * m, hipMemcpyDeviceToHost, s));
        PBD_HIP(hipStreamSynchronize(s));
    });
}

int pbd_ccd_resolve(pbd_ctx* ctx, float* p, const float* w, float th, int passes) {
    return guard([&] {
        PBD_CHECK(ctx && p, PBD_ERR_ARG, "null argument");
        PBD_CHECK(ctx->ccd_built, PBD_ERR_STATE, "call pbd_ccd_detect first");
        PBD_CHECK(passes >= 0, PBD_ERR_ARG, "passes must be >= 0");
        set_device(ctx);
        hipStream_t s = ctx->stream;
        int nv = ctx->snv;
        std::vector<float4> P(nv), X;
        ctx->sx.download(X, s);
        PBD_HIP(hipStreamSynchronize(s));
        for (int i = 0; i < nv; i++) {
            float wi = w ? w[i] : 1.0f;
            P[i] = make_float4(p[3 * i], p[3 * i + 1], p[3 * i + 2], wi);
            X[i].w = wi;
        }
        ctx->sp4.upload(P, s);
        ctx->sx.upload(X, s);
        ctx->sbp.resolve(s, ctx->sx.ptr, ctx->sp4.ptr, ctx->stris.ptr, nv, th, passes, 1e-6f);
        ctx->sp4.download(P, s);
        PBD_HIP(hipStreamSynchronize(s));
        for (int i = 0; i < nv; i++) {
            p[3 * i] = P[i].x;
            p[3 * i + 1] = P[i].y;
            p[3 * i + 2] = P[i].z;
        }
    });
}

int pbd_set_collision(pbd_ctx* ctx, const pbd_collision_params* cp) {
    return guard([&] {
        PBD_CHECK(ctx && cp, PBD_ERR_ARG, "null argument");
        PBD_CHECK(ctx->inited, PBD_ERR_STATE, "call pbd_init first");
        set_device(ctx);
        drop_graphs(ctx);
        if (!cp->enabled) {
            ctx->coll = *cp;
            return;
        }
        PBD_CHECK(cp->thickness >= 0.0f && cp->resolve_passes >= 0 && cp->min_normal2 >= 0.0f &&
                      (cp->broadphase == 0 || cp->broadphase == 1) && (cp->edge_edge == 0 || cp->edge_edge == 1),
                  PBD_ERR_ARG, "bad collision parameters");
        pbd_hash_desc d = cp->grid;
        if (cp->cell_scale > 0.0f) {
            double sum = 0.0;
            long long n = 0;
            for (auto& k : ctx->cloths)
                for (int p = 0; p < k.np(); p++)
                    if (k.prim_list[2 * p + 1] == 2) {
                        sum += k.rest[p];
                        n++;
                    }
            PBD_CHECK(n > 0, PBD_ERR_STATE, "cell_scale needs distance prims for the mean edge length");
            float cell = (float)(cp->cell_scale * (sum / n));
            d.cell[0] = d.cell[1] = d.cell[2] = cell;
        }
        ctx->ccfg = make_hash_cfg(d);
        std::vector<int> gt;
        for (auto& k : ctx->cloths)
            for (int i : k.tris) gt.push_back(i + k.vbase);
        ctx->gtris.upload(gt, ctx->stream);
        ctx->gnt = (int)gt.size() / 3;
        if (cp->edge_edge) {
            std::vector<int> te;
            std::vector<int2> ed;
            build_edge_table(gt.data(), ctx->gnt, te, ed);
            ctx->cbp.set_edges(ctx->stream, te, ed);
        }
        ctx->coll = *cp;
        replan_if_needed(ctx);
        PBD_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int pbd_last_contacts(pbd_ctx* ctx, long long* nc, long long* np) {
    return guard([&] {
        PBD_CHECK(ctx, PBD_ERR_ARG, "null context");
        if (nc) *nc = ctx->cbp.ncontacts;
        if (np) *np = ctx->cbp.npairs;
    });
}

int pbd_ws_init(pbd_ctx* ctx, uint64_t seed) {
    return guard([&] {
        PBD_CHECK(ctx, PBD_ERR_ARG, "null context");
        set_device(ctx);
        int n = mlp_nparam();
        std::vector<float> w(n, 0.0f);
        uint64_t st = seed;
        auto fill = [&](int off, int rows, int cols, float scale) {
            float a = scale * std::sqrt(6.0f / (float)(rows + cols));
            for (int i = 0; i < rows * cols; i++) w[off + i] = a * uniform_pm1(&st);
        };
        // Glorot-uniform weights, zero biases; a small output layer so the
        // untrained prediction starts near zero
        fill(0, 64, 22, 1.0f);
        fill(64 * 22 + 64, 64, 64, 1.0f);
        fill(64 * 22 + 64 + 4096 + 64, 3, 64, 0.1f);
        auto& W = ctx->ws;
        W.prm.upload(w, ctx->stream);
        W.m.upload(std::vector<float>(n, 0.0f), ctx->stream);
        W.v.upload(std::vector<float>(n, 0.0f), ctx->stream);
        W.grad.alloc(n + 1);
        W.W.alloc(mlp_wbytes());
        mlp_pack(ctx->stream, W.prm.ptr, W.W.ptr);
        W.adam_step = 0;
        W.params_ready = true;
        PBD_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int pbd_ws_num_params(int* n) {
    return guard([&] {
        PBD_CHECK(n, PBD_ERR_ARG, "null output");
        *n = mlp_nparam();
    });
}

int pbd_ws_get_params(pbd_ctx* ctx, float* out) {
    return guard([&] {
        PBD_CHECK(ctx && out, PBD_ERR_ARG, "null argument");
        ws_require_params(ctx);
        set_device(ctx);
        PBD_HIP(hipMemcpyAsync(out, ctx->ws.prm.ptr, sizeof(float) * mlp_nparam(), hipMemcpyDeviceToHost, ctx->stream));
        PBD_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int pbd_ws_set_params(pbd_ctx* ctx, const float* in) {
    return guard([&] {
        PBD_CHECK(ctx && in, PBD_ERR_ARG, "null argument");
        set_device(ctx);
        int n = mlp_nparam();
        auto& W = ctx->ws;
        W.prm.upload(std::vector<float>(in, in + n), ctx->stream);
        if (!W.params_ready) {
            W.m.upload(std::vector<float>(n, 0.0f), ctx->stream);
            W.v.upload(std::vector<float>(n, 0.0f), ctx->stream);
            W.grad.alloc(n + 1);
            W.W.alloc(mlp_wbytes());
        }
        mlp_pack(ctx->stream, W.prm.ptr, W.W.ptr);
        W.params_ready = true;
        PBD_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int pbd_ws_forward(pbd_ctx* ctx, const float* X, int n, float* Y) {
    return guard([&] {
        PBD_CHECK(ctx && (X || !n) && (Y || !n) && n >= 0, PBD_ERR_ARG, "bad argument");
        ws_require_params(ctx);
        set_device(ctx);
        DBuf<float> dx, dy;
        dx.upload(std::vector<float>(X, X + 22ll * n), ctx->stream);
        dy.alloc(3ll * std::max(n, 1));
        mlp_forward(ctx->stream, ctx->ws.W.ptr, dx.ptr, n, dy.ptr);
        if (n) PBD_HIP(hipMemcpyAsync(Y, dy.ptr, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, ctx->stream));
        PBD_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int pbd_ws_set_samples(pbd_ctx* ctx, const float* X, const float* Y, long long n) {
    return guard([&] {
        PBD_CHECK(ctx && (X || !n) && (Y || !n) && n >= 0 && n < (1ll << 31), PBD_ERR_ARG, "bad argument");
        set_device(ctx);
        ctx->ws.X.upload(std::vector<float>(X, X + 22 * n), ctx->stream);
        ctx->ws.Y.upload(std::vector<float>(Y, Y + 3 * n), ctx->stream);
        ctx->ws.ns = n;
        PBD_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int pbd_ws_num_samples(pbd_ctx* ctx, long long* n) {
    return guard([&] {
        PBD_CHECK(ctx && n, PBD_ERR_ARG, "null argument");
        *n = ctx->ws.ns;
    });
}

int pbd_ws_get_samples(pbd_ctx* ctx, long long first, long long n, float* X, float* Y) {
    return guard([&] {
        PBD_CHECK(ctx && first >= 0 && n >= 0 && first + n <= ctx->ws.ns, PBD_ERR_ARG, "sample range out of bounds");
        set_device(ctx);
        if (n && X) PBD_HIP(hipMemcpyAsync(X, ctx->ws.X.ptr + 22 * first, sizeof(float) * 22 * n, hipMemcpyDeviceToHost, ctx->stream));
        if (n && Y) PBD_HIP(hipMemcpyAsync(Y, ctx->ws.Y.ptr + 3 * first, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, ctx->stream));
        PBD_HIP(hipStreamSynchronize(ctx->stream));
    });
}

// Runs `substeps` plain substeps of every cloth with `sweeps` Gauss-Seidel
// sweeps each and records, per vertex, the features at the pre-projection
// state and the correction those sweeps produced.
int pbd_ws_collect(pbd_ctx* ctx, int sweeps, int substeps) {
    return guard([&] {
        PBD_CHECK(ctx && ctx->inited, PBD_ERR_STATE, "call pbd_init first");
        PBD_CHECK(sweeps >= 1 && substeps >= 0, PBD_ERR_ARG, "bad sweep / substep count");
        ws_require_params(ctx);
        PBD_CHECK(!ctx->ws.samples_normalized, PBD_ERR_STATE, "samples were standardised by training; collect before pbd_ws_train");
        set_device(ctx);
        ws_build_nbr(ctx);
        hipStream_t s = ctx->stream;
        auto& W = ctx->ws;
        long long need = W.ns + (long long)substeps * ctx->nv_total;
        PBD_CHECK(need < (1ll << 31), PBD_ERR_LIMIT, "too many samples");
        if ((long long)W.X.n < 22 * need) {
            DBuf<float> nx, ny;
            nx.alloc(22 * need);
            ny.alloc(3 * need);
            if (W.ns) {
                PBD_HIP(hipMemcpyAsync(nx.ptr, W.X.ptr, sizeof(float) * 22 * W.ns, hipMemcpyDeviceToDevice, s));
                PBD_HIP(hipMemcpyAsync(ny.ptr, W.Y.ptr, sizeof(float) * 3 * W.ns, hipMemcpyDeviceToDevice, s));
            }
            std::swap(W.X.ptr, nx.ptr);
            std::swap(W.X.n, nx.n);
            std::swap(W.Y.ptr, ny.ptr);
            std::swap(W.Y.n, ny.n);
        }
        // stiffness tables for `sweeps` sweeps while recording; the solver's
        // own budget is restored afterwards
        const int budget = ctx->sp.iterations;
        auto use_budget = [&](int it) {
            if (ctx->sp.iterations == it) return;
            ctx->sp.iterations = it;
            fill_ktabs(ctx);
            fill_cloth_dev(ctx);
        };
        use_budget(sweeps);
        struct Restore {
            std::function<void()> f;
            ~Restore() {
                try { f(); } catch (...) {}
            }
        } restore{[&] { use_budget(budget); }};
        float dt = ctx->sp.dt;
        for (int k = 0; k < substeps; k++) {
            advect_all(ctx, dt);
            float4* p = ctx->pbuf[ctx->pcur].ptr;
            ws_infer(s, W.W.ptr, ctx->cl.ptr, ctx->vcloth.ptr, ctx->x.ptr, p, nullptr, 0.0f, W.nbr.ptr, W.nbr_rest.ptr,
                     ctx->nv_total, W.delta.ptr, W.X.ptr + 22 * W.ns, W.norm.ptr, W.feat.ptr, W.hs.ptr, 0.0f);
            PBD_HIP(hipMemcpyAsync(W.psave.ptr, p, sizeof(float4) * ctx->nv_total, hipMemcpyDeviceToDevice, s));
            double tol = ctx->sp.tol;
            ctx->sp.tol = -1.0;
            if (ctx->coll.enabled) {
                ccd_sweeps(ctx, sweeps, dt, nullptr);  // the target includes the contact response
            } else {
                project(ctx, sweeps, false, dt);
                consolidate(ctx);
            }
            ctx->sp.tol = tol;
            ws_target(s, ctx->cl.ptr, ctx->vcloth.ptr, ctx->pbuf[ctx->pcur].ptr, W.psave.ptr, ctx->nv_total,
                      W.Y.ptr + 3 * W.ns);
            integrate_all(ctx, dt, false);
            W.ns += ctx->nv_total;
        }
        PBD_HIP(hipStreamSynchronize(s));
    });
}

// Gradient of the mean loss over samples [first, first+n) (tests).
int pbd_ws_gradient(pbd_ctx* ctx, long long first, int n, float* grad, double* loss) {
    return guard([&] {
        PBD_CHECK(ctx && grad && n > 0 && first >= 0 && first + n <= ctx->ws.ns, PBD_ERR_ARG, "bad sample range");
        ws_require_params(ctx);
        set_device(ctx);
        hipStream_t s = ctx->stream;
        std::vector<int> idx(n);
        for (int i = 0; i < n; i++) idx[i] = (int)(first + i);
        auto& W = ctx->ws;
        W.perm.upload(idx, s);
        W.partial.alloc((size_t)(((n + 15) / 16 + 2) * (mlp_nparam() + 1)));
        mlp_grad(s, W.W.ptr, W.X.ptr, W.Y.ptr, W.perm.ptr, n, 1.0f / n, W.partial.ptr, W.grad.ptr);
        std::vector<float> g;
        W.grad.download(g, s);
        PBD_HIP(hipStreamSynchronize(s));
        std::copy(g.begin(), g.begin() + mlp_nparam(), grad);
        if (loss) *loss = g[mlp_nparam()];
    });
}

int pbd_ws_adam_step(pbd_ctx* ctx, const float* grad, float lr) {
    return guard([&] {
        PBD_CHECK(ctx && grad, PBD_ERR_ARG, "null argument");
        ws_require_params(ctx);
        set_device(ctx);
        auto& W = ctx->ws;
        std::vector<float> g(grad, grad + mlp_nparam());
        g.push_back(0.0f);
        W.grad.upload(g, ctx->stream);
        mlp_adam(ctx->stream, W.prm.ptr, W.grad.ptr, W.m.ptr, W.v.ptr, lr, 0.9f, 0.999f, 1e-8f, ++W.adam_step);
        mlp_pack(ctx->stream, W.prm.ptr, W.W.ptr);
        PBD_HIP(hipStreamSynchronize(ctx->stream));
    });
}

// Minibatch Adam over the collected samples, reshuffled every epoch.
int pbd_ws_train(pbd_ctx* ctx, int epochs, int batch, float lr, uint64_t seed, double* last_loss) {
    return guard([&] {
        PBD_CHECK(ctx && epochs >= 0 && batch >= 16, PBD_ERR_ARG, "bad training arguments");
        ws_require_params(ctx);
        PBD_CHECK(ctx->ws.ns > 0, PBD_ERR_STATE, "no samples");
        set_device(ctx);
        hipStream_t s = ctx->stream;
        auto& W = ctx->ws;
        long long n = W.ns;
        if (!W.samples_normalized) {
            // grouped scales: one RMS for the displacement (3), one shared
            // by the 18 strain components (keeps the scaling rotation
            // equivariant), 1 for the pin flag, one RMS for the 3 targets;
            // no centring
            std::vector<float> X, Y;
            W.X.download(X, s);
            W.Y.download(Y, s);
            PBD_HIP(hipStreamSynchronize(s));
            double sd = 0.0, ss = 0.0, sy = 0.0;
            for (long long i = 0; i < n; i++) {
                for (int k = 0; k < 3; k++) sd += (double)X[22 * i + k] * X[22 * i + k];
                for (int k = 3; k < 21; k++) ss += (double)X[22 * i + k] * X[22 * i + k];
                for (int k = 0; k < 3; k++) sy += (double)Y[3 * i + k] * Y[3 * i + k];
            }
            auto rms = [](double v, double cnt) { double r = std::sqrt(v / cnt); return r > 1e-20 ? (float)r : 1.0f; };
            W.norm_h.assign(47, 0.0f);
            for (int k = 0; k < 3; k++) W.norm_h[22 + k] = rms(sd, 3.0 * n);
            for (int k = 3; k < 21; k++) W.norm_h[22 + k] = rms(ss, 18.0 * n);
            W.norm_h[22 + 21] = 1.0f;
            for (int k = 0; k < 3; k++) W.norm_h[44 + k] = rms(sy, 3.0 * n);
            W.norm.upload(W.norm_h, s);
            ws_normalize(s, W.X.ptr, W.Y.ptr, n, W.norm.ptr);
            W.samples_normalized = true;
        }
        std::vector<int> perm(n);
        for (long long i = 0; i < n; i++) perm[i] = (int)i;
        uint64_t st = seed;
        W.partial.alloc((size_t)(((batch + 15) / 16 + 2) * (mlp_nparam() + 1)));
        double loss_sum = 0.0;
        long long loss_n = 0;
        std::vector<float> g(1);
        for (int e = 0; e < epochs; e++) {
            for (long long i = n - 1; i > 0; i--) {
                long long j = (long long)(splitmix64(&st) % (uint64_t)(i + 1));
                std::swap(perm[i], perm[j]);
            }
            W.perm.upload(perm, s);
            loss_sum = 0.0;
            loss_n = 0;
            for (long long b0 = 0; b0 + batch <= n || (b0 == 0 && n < batch); b0 += batch) {
                int m = (int)std::min<long long>(batch, n - b0);
                mlp_grad(s, W.W.ptr, W.X.ptr, W.Y.ptr, W.perm.ptr + b0, m, 1.0f / m, W.partial.ptr, W.grad.ptr);
                mlp_adam(s, W.prm.ptr, W.grad.ptr, W.m.ptr, W.v.ptr, lr, 0.9f, 0.999f, 1e-8f, ++W.adam_step);
                mlp_pack(s, W.prm.ptr, W.W.ptr);
                if (e == epochs - 1) {
                    float lv = 0.0f;
                    PBD_HIP(hipMemcpyAsync(&lv, W.grad.ptr + mlp_nparam(), sizeof(float), hipMemcpyDeviceToHost, s));
                    PBD_HIP(hipStreamSynchronize(s));
                    loss_sum += lv;
                    loss_n++;
                }
            }
        }
        PBD_HIP(hipStreamSynchronize(s));
        if (last_loss) *last_loss = loss_n ? loss_sum / loss_n : 0.0;
    });
}

int pbd_ws_get_normalization(pbd_ctx* ctx, float* norm47) {
    return guard([&] {
        PBD_CHECK(ctx && norm47, PBD_ERR_ARG, "null argument");
        PBD_CHECK(ctx->ws.norm_h.size() == 47, PBD_ERR_STATE, "no normalisation yet");
        std::copy(ctx->ws.norm_h.begin(), ctx->ws.norm_h.end(), norm47);
    });
}

int pbd_ws_set_normalization(pbd_ctx* ctx, const float* norm47) {
    return guard([&] {
        PBD_CHECK(ctx && norm47, PBD_ERR_ARG, "null argument");
        for (int k = 22; k < 47; k++) PBD_CHECK(norm47[k] > 0.0f, PBD_ERR_ARG, "scales must be positive");
        set_device(ctx);
        ctx->ws.norm_h.assign(norm47, norm47 + 47);
        ctx->ws.norm.upload(ctx->ws.norm_h, ctx->stream);
        PBD_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int pbd_cloth_set_tolerance(pbd_ctx* ctx, int cloth, double tol) {
    return guard([&] {
        Cloth& c = cloth_of(ctx, cloth);
        c.tol = tol;
        if (ctx->inited) {
            set_device(ctx);
            fill_cloth_dev(ctx);
        }
    });
}

int pbd_ws_enable(pbd_ctx* ctx, int on) {
    return guard([&] {
        PBD_CHECK(ctx && ctx->inited, PBD_ERR_STATE, "call pbd_init first");
        if (on) ws_require_params(ctx);
        ctx->ws.enabled = on != 0;
        drop_graphs(ctx);
    });
}

int pbd_ws_last_used(pbd_ctx* ctx, int cloth, int* used) {
    return guard([&] {
        cloth_of(ctx, cloth);
        PBD_CHECK(used, PBD_ERR_ARG, "null output");
        set_device(ctx);
        ws_read_used(ctx);
        *used = cloth < (int)ctx->ws.used.size() ? ctx->ws.used[cloth] : 0;
    });
}

int pbd_ws_last_delta(pbd_ctx* ctx, int cloth, float* xyz) {
    return guard([&] {
        Cloth& k = cloth_of(ctx, cloth);
        PBD_CHECK(xyz, PBD_ERR_ARG, "null output");
        PBD_CHECK(ctx->ws.nbr_built && ctx->ws.delta.ptr, PBD_ERR_STATE, "no prediction made yet");
        set_device(ctx);
        std::vector<float4> d(k.nv);
        PBD_HIP(hipMemcpyAsync(d.data(), ctx->ws.delta.ptr + k.vbase, sizeof(float4) * k.nv, hipMemcpyDeviceToHost,
                               ctx->stream));
        PBD_HIP(hipStreamSynchronize(ctx->stream));
        for (int i = 0; i < k.nv; i++) {
            xyz[3 * i] = d[i].x;
            xyz[3 * i + 1] = d[i].y;
            xyz[3 * i + 2] = d[i].z;
        }
    });
}

int pbd_profile_enable(pbd_ctx* ctx, int on) {
    return guard([&] {
        PBD_CHECK(ctx, PBD_ERR_ARG, "null context");
        ctx->prof_on = on != 0;
    });
}

int pbd_profile_read(pbd_ctx* ctx, pbd_profile* out, int reset) {
    return guard([&] {
        PBD_CHECK(ctx && out, PBD_ERR_ARG, "null argument");
        set_device(ctx);
        *out = ctx->prof;
        const Broadphase& b = ctx->cbp;
        out->ccd_ee_raw_pairs = b.st_ee_raw;
        out->ccd_ee_pairs = b.st_ee_pairs;
        out->ccd_ee_contacts = b.st_ee_contacts;
        out->ccd_detections = b.st_dets;
        out->ccd_triangles = b.st_tris;
        out->ccd_cell_entries = b.st_entries;
        out->ccd_raw_pairs = b.st_raw;
        out->ccd_pairs = b.st_pairs;
        out->ccd_contacts = b.st_contacts;
        out->ccd_resolved = b.st_resolved;
        out->queue_wait_timeouts = 0;
        out->queue_deciding_launches = 0;
        if (ctx->q_timeouts.n) {
            int t[2] = {0, 0};
            PBD_HIP(hipMemcpyAsync(t, ctx->q_timeouts.ptr, 2 * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
            PBD_HIP(hipStreamSynchronize(ctx->stream));
            out->queue_wait_timeouts = t[0];
            out->queue_deciding_launches = t[1];
        }
        if (reset) {
            ctx->prof = pbd_profile{};
            ctx->cbp.st_dets = ctx->cbp.st_tris = ctx->cbp.st_entries = 0;
            ctx->cbp.st_raw = ctx->cbp.st_pairs = ctx->cbp.st_contacts = ctx->cbp.st_resolved = 0;
            ctx->cbp.st_ee_raw = ctx->cbp.st_ee_pairs = ctx->cbp.st_ee_contacts = 0;
        }
    });
}

int pbd_bench_copy(pbd_ctx* ctx, long long bytes, int reps, double* ms) {
    return guard([&] {
        PBD_CHECK(ctx && ms, PBD_ERR_ARG, "null argument");
        PBD_CHECK(bytes >= 16 && reps >= 1, PBD_ERR_ARG, "bad copy size");
        set_device(ctx);
        size_t n = (size_t)bytes / 16;
        DBuf<float4> a, b;
        a.alloc(n);
        b.alloc(n);
        PBD_HIP(hipMemsetAsync(a.ptr, 0, n * 16, ctx->stream));
        launch_copy(ctx->stream, a.ptr, b.ptr, n);  // warm-up
        hipEvent_t e0 = next_event(ctx), e1 = next_event(ctx);
        ctx->ev_used -= 2;
        PBD_HIP(hipEventRecord(e0, ctx->stream));
        for (int r = 0; r < reps; r++) launch_copy(ctx->stream, a.ptr, b.ptr, n);
        PBD_HIP(hipEventRecord(e1, ctx->stream));
        PBD_HIP(hipStreamSynchronize(ctx->stream));
        float t = 0.0f;
        PBD_HIP(hipEventElapsedTime(&t, e0, e1));
        *ms = (double)t / reps;
    });
}

int pbd_cloth_get_state(pbd_ctx* ctx, int id, int which, float* xyz) {
    return guard([&] {
        Cloth& c = cloth_of(ctx, id);
        PBD_CHECK(ctx->inited, PBD_ERR_STATE, "call pbd_init first");
        PBD_CHECK(xyz, PBD_ERR_ARG, "null output");
        PBD_CHECK(which >= PBD_STATE_X && which <= PBD_STATE_P, PBD_ERR_ARG, "unknown state buffer");
        set_device(ctx);
        if (which == PBD_STATE_P) consolidate(ctx);
        DBuf<float4>& b = which == PBD_STATE_X ? ctx->x : which == PBD_STATE_V ? ctx->v : ctx->pbuf[ctx->pcur];
        std::vector<float4> h(c.nv);
        PBD_HIP(hipMemcpyAsync(h.data(), b.ptr + c.vbase, sizeof(float4) * c.nv, hipMemcpyDeviceToHost, ctx->stream));
        PBD_HIP(hipStreamSynchronize(ctx->stream));
        for (int i = 0; i < c.nv; i++) {
            xyz[3 * i] = h[i].x;
            xyz[3 * i + 1] = h[i].y;
            xyz[3 * i + 2] = h[i].z;
        }
    });
}

int pbd_cloth_set_state(pbd_ctx* ctx, int id, int which, const float* xyz) {
    return guard([&] {
        Cloth& c = cloth_of(ctx, id);
        PBD_CHECK(ctx->inited, PBD_ERR_STATE, "call pbd_init first");
        PBD_CHECK(xyz, PBD_ERR_ARG, "null input");
        PBD_CHECK(which >= PBD_STATE_X && which <= PBD_STATE_P, PBD_ERR_ARG, "unknown state buffer");
        set_device(ctx);
        if (which == PBD_STATE_P) consolidate(ctx);  // the other cloths' p must be in place first
        DBuf<float4>& b = which == PBD_STATE_X ? ctx->x : which == PBD_STATE_V ? ctx->v : ctx->pbuf[ctx->pcur];
        std::vector<float4> h(c.nv);
        for (int i = 0; i < c.nv; i++)
            h[i] = make_float4(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], which == PBD_STATE_V ? 0.0f : c.w[i]);
        PBD_HIP(hipMemcpyAsync(b.ptr + c.vbase, h.data(), sizeof(float4) * c.nv, hipMemcpyHostToDevice, ctx->stream));
        PBD_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int pbd_cloth_residual(pbd_ctx* ctx, int id, double* max_rel, double* rms) {
    return guard([&] {
        cloth_of(ctx, id);
        PBD_CHECK(ctx->inited, PBD_ERR_STATE, "call pbd_init first");
        set_device(ctx);
        hipStream_t s = ctx->stream;
        int ncl = (int)ctx->cloths.size();
        consolidate(ctx);
        launch_fill_u64(s, ctx->rmax.ptr, ncl, 0ull);
        PBD_HIP(hipMemsetAsync(ctx->rsum.ptr, 0, sizeof(double) * ncl, s));
        launch_residual(s, ctx->rdist.ptr, (int)ctx->rdist.n, ctx->pbuf[ctx->pcur].ptr, ctx->rmax.ptr, ctx->rsum.ptr,
                        nullptr);
        unsigned long long m = 0;
        double sum = 0.0;
        PBD_HIP(hipMemcpyAsync(&m, ctx->rmax.ptr + id, 8, hipMemcpyDeviceToHost, s));
        PBD_HIP(hipMemcpyAsync(&sum, ctx->rsum.ptr + id, 8, hipMemcpyDeviceToHost, s));
        PBD_HIP(hipStreamSynchronize(s));
        double mv;
        memcpy(&mv, &m, 8);
        if (max_rel) *max_rel = mv;
        int ne = ctx->cloths[id].ne;
        if (rms) *rms = ne > 0 ? std::sqrt(sum / ne) : 0.0;
    });
}

int pbd_residual_all(pbd_ctx* ctx, double* max_rel, double* rms) {
    return guard([&] {
        PBD_CHECK(ctx && ctx->inited, PBD_ERR_STATE, "call pbd_init first");
        set_device(ctx);
        hipStream_t s = ctx->stream;
        int ncl = (int)ctx->cloths.size();
        consolidate(ctx);
        launch_fill_u64(s, ctx->rmax.ptr, ncl, 0ull);
        PBD_HIP(hipMemsetAsync(ctx->rsum.ptr, 0, sizeof(double) * ncl, s));
        launch_residual(s, ctx->rdist.ptr, (int)ctx->rdist.n, ctx->pbuf[ctx->pcur].ptr, ctx->rmax.ptr, ctx->rsum.ptr,
                        nullptr);
        std::vector<unsigned long long> m;
        std::vector<double> sum;
        ctx->rmax.download(m, s);
        ctx->rsum.download(sum, s);
        PBD_HIP(hipStreamSynchronize(s));
        for (int i = 0; i < ncl; i++) {
            double mv;
            memcpy(&mv, &m[i], 8);
            if (max_rel) max_rel[i] = mv;
            int ne = ctx->cloths[i].ne;
            if (rms) rms[i] = ne > 0 ? std::sqrt(sum[i] / ne) : 0.0;
        }
    });
}

// Whole-context state (every cloth, packed vertex order) to / from host
// arrays of 3 * total-vertices floats: an exact binary checkpoint of x and v.
int pbd_state_save(pbd_ctx* ctx, float* x, float* v) {
    return guard([&] {
        PBD_CHECK(ctx && ctx->inited, PBD_ERR_STATE, "call pbd_init first");
        set_device(ctx);
        std::vector<float4> X, V;
        ctx->x.download(X, ctx->stream);
        ctx->v.download(V, ctx->stream);
        PBD_HIP(hipStreamSynchronize(ctx->stream));
        for (int i = 0; i < ctx->nv_total; i++) {
            if (x) { x[3 * i] = X[i].x; x[3 * i + 1] = X[i].y; x[3 * i + 2] = X[i].z; }
            if (v) { v[3 * i] = V[i].x; v[3 * i + 1] = V[i].y; v[3 * i + 2] = V[i].z; }
        }
    });
}

int pbd_state_load(pbd_ctx* ctx, const float* x, const float* v) {
    return guard([&] {
        PBD_CHECK(ctx && ctx->inited, PBD_ERR_STATE, "call pbd_init first");
        PBD_CHECK(x && v, PBD_ERR_ARG, "null state");
        set_device(ctx);
        std::vector<float4> X, V(ctx->nv_total);
        ctx->x.download(X, ctx->stream);
        PBD_HIP(hipStreamSynchronize(ctx->stream));
        for (int i = 0; i < ctx->nv_total; i++) {
            X[i] = make_float4(x[3 * i], x[3 * i + 1], x[3 * i + 2], X[i].w);
            V[i] = make_float4(v[3 * i], v[3 * i + 1], v[3 * i + 2], 0.0f);
        }
        ctx->x.upload(X, ctx->stream);
        ctx->v.upload(V, ctx->stream);
        PBD_HIP(hipMemcpyAsync(ctx->pbuf[ctx->pcur].ptr, ctx->x.ptr, sizeof(float4) * ctx->nv_total,
                               hipMemcpyDeviceToDevice, ctx->stream));
        ctx->pfinal_valid = false;
        PBD_HIP(hipStreamSynchronize(ctx->stream));
    });
}

// Device-side copy of x and v (and restore), for A/B runs from one state.
int pbd_snapshot(pbd_ctx* ctx, int restore) {
    return guard([&] {
        PBD_CHECK(ctx && ctx->inited, PBD_ERR_STATE, "call pbd_init first");
        set_device(ctx);
        hipStream_t s = ctx->stream;
        size_t b = sizeof(float4) * ctx->nv_total;
        if (!restore) {
            ctx->snap_x.alloc(ctx->nv_total);
            ctx->snap_v.alloc(ctx->nv_total);
            PBD_HIP(hipMemcpyAsync(ctx->snap_x.ptr, ctx->x.ptr, b, hipMemcpyDeviceToDevice, s));
            PBD_HIP(hipMemcpyAsync(ctx->snap_v.ptr, ctx->v.ptr, b, hipMemcpyDeviceToDevice, s));
            ctx->snap_valid = true;
        } else {
            PBD_CHECK(ctx->snap_valid, PBD_ERR_STATE, "no snapshot taken");
            PBD_HIP(hipMemcpyAsync(ctx->x.ptr, ctx->snap_x.ptr, b, hipMemcpyDeviceToDevice, s));
            PBD_HIP(hipMemcpyAsync(ctx->v.ptr, ctx->snap_v.ptr, b, hipMemcpyDeviceToDevice, s));
            PBD_HIP(hipMemcpyAsync(ctx->pbuf[ctx->pcur].ptr, ctx->snap_x.ptr, b, hipMemcpyDeviceToDevice, s));
            ctx->pfinal_valid = false;
        }
        PBD_HIP(hipStreamSynchronize(s));
    });
}

int pbd_autotune(pbd_ctx* ctx, int trial_substeps, pbd_solver_params* best, int* best_table_bits,
                 double* best_ms_per_substep, int* trials) {
    return guard([&] {
        PBD_CHECK(ctx && ctx->inited, PBD_ERR_STATE, "call pbd_init first");
        PBD_CHECK(trial_substeps >= 1, PBD_ERR_ARG, "trial_substeps must be >= 1");
        set_device(ctx);
        hipStream_t s = ctx->stream;
        const size_t nv = (size_t)ctx->nv_total;
        DBuf<float4> sx, sv;
        sx.alloc(nv);
        sv.alloc(nv);
        launch_copy(s, ctx->x.ptr, sx.ptr, nv);
        launch_copy(s, ctx->v.ptr, sv.ptr, nv);
        auto restore = [&] {
            launch_copy(s, sx.ptr, ctx->x.ptr, nv);
            launch_copy(s, sv.ptr, ctx->v.ptr, nv);
            launch_copy(s, sx.ptr, ctx->pbuf[ctx->pcur].ptr, nv);
            ctx->pfinal_valid = false;
        };
        const pbd_solver_params base = ctx->sp;
        const HashCfg hbase = ctx->ccfg;
        // tile sizes: the whole LDS, half of it (two workgroups per CU), and two fixed caps
        const int half = (ctx->lds_max / 2 - TILE_STATIC_LDS) / 16;
        std::vector<int> threads = {256, 512, 1024}, sweeps, maxv = {0, 4096, 2048}, bits = {-1};
        if (half > 4096 && half < effective_max_load(ctx)) maxv.insert(maxv.begin() + 1, half);
        if (ctx->coll.enabled) {
            sweeps = {base.tile_sweeps};  // contacts force one sweep per launch
            if (hbase.mode == 1) {
                bits.clear();
                for (int d = -2; d <= 2; d++)
                    if (hbase.table_bits + d >= 8 && hbase.table_bits + d <= 26) bits.push_back(hbase.table_bits + d);
            }
        } else if (has_tolerance(ctx)) {
            sweeps = {std::max(1, base.check_every)};  // chunk = check interval keeps the stopping points
        } else {
            for (int t : {1, 2, 4, 5})
                if (t <= base.iterations) sweeps.push_back(t);
        }
        hipEvent_t e0 = nullptr, e1 = nullptr;
        PBD_HIP(hipEventCreate(&e0));
        PBD_HIP(hipEventCreate(&e1));
        double best_t = 1e300;
        pbd_solver_params bp = base;
        int bb = -1, n = 0;
        try {
            for (int hb : bits) {
                if (hb >= 0) {
                    ctx->ccfg.table_bits = hb;
                    ctx->ccfg.ncells = 1 << hb;
                }
                for (int ts : sweeps)
                    for (int mv : maxv)
                        for (int th : threads) {
                            ctx->sp = base;
                            ctx->sp.tile_sweeps = ts;
                            ctx->sp.tile_max_verts = mv;
                            ctx->sp.block_threads = th;
                            drop_graphs(ctx);
                            try {
                                replan_if_needed(ctx);
                            } catch (const Error& e) {
                                if (e.code != PBD_ERR_LIMIT) throw;
                                continue;  // no tile plan fits this candidate
                            }
                            restore();
                            run_substeps(ctx, 1);  // untimed: capture, first-touch
                            restore();
                            PBD_HIP(hipEventRecord(e0, s));
                            run_substeps(ctx, trial_substeps);
                            PBD_HIP(hipEventRecord(e1, s));
                            PBD_HIP(hipEventSynchronize(e1));
                            float ms = 0.0f;
                            PBD_HIP(hipEventElapsedTime(&ms, e0, e1));
                            n++;
                            if (ms < best_t) {
                                best_t = ms;
                                bp = ctx->sp;
                                bb = hb;
                            }
                        }
            }
            // then the other way of stepping single-tile cloths, at the best setting
            if (!ctx->coll.enabled) {
                ctx->sp = bp;
                ctx->sp.split_whole = bp.split_whole ? 0 : 1;
                drop_graphs(ctx);
                replan_if_needed(ctx);
                restore();
                run_substeps(ctx, 1);
                restore();
                PBD_HIP(hipEventRecord(e0, s));
                run_substeps(ctx, trial_substeps);
                PBD_HIP(hipEventRecord(e1, s));
                PBD_HIP(hipEventSynchronize(e1));
                float ms = 0.0f;
                PBD_HIP(hipEventElapsedTime(&ms, e0, e1));
                n++;
                if (ms < best_t) {
                    best_t = ms;
                    bp = ctx->sp;
                }
            }
        } catch (...) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            ctx->sp = base;
            ctx->ccfg = hbase;
            drop_graphs(ctx);
            throw;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        ctx->sp = bp;
        if (bb >= 0) {
            ctx->ccfg.table_bits = bb;
            ctx->ccfg.ncells = 1 << bb;
        }
        drop_graphs(ctx);
        replan_if_needed(ctx);
        restore();
        PBD_HIP(hipStreamSynchronize(s));
        if (best) *best = bp;
        if (best_table_bits) *best_table_bits = bb;
        if (best_ms_per_substep) *best_ms_per_substep = best_t / trial_substeps;
        if (trials) *trials = n;
    });
}

int pbd_num_vertices(pbd_ctx* ctx, int* nv) {
    return guard([&] {
        PBD_CHECK(ctx && nv && ctx->inited, PBD_ERR_STATE, "call pbd_init first");
        *nv = ctx->nv_total;
    });
}

int pbd_cloth_last_sweeps(pbd_ctx* ctx, int id, int* sweeps) {
    return guard([&] {
        Cloth& c = cloth_of(ctx, id);
        PBD_CHECK(sweeps, PBD_ERR_ARG, "null output");
        set_device(ctx);
        read_sweeps(ctx);
        *sweeps = c.last_sweeps;
    });
}

}  // extern "C"
