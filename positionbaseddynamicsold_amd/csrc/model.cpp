// Host-side model build: grid/mesh generation, constraint prims, adjacency,
// the reference text loader, and the LDS tile planner.
//
// Reference paths are relative to PositionBasedDynamicsCUDA/PositionBasedDynamicsCUDA/.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <fstream>
#include <sstream>
#include <thread>
#include <unordered_map>

#include "pbd_internal.hpp"
#include "pbd_math.hpp"

namespace pbd {

// splitmix64 (Steele, Lea, Flood 2014); uniform_pm1 maps the top 24 bits to [-1,1).
uint64_t splitmix64(uint64_t* s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

float uniform_pm1(uint64_t* s) {
    uint64_t r = splitmix64(s) >> 40;
    float u = (float)r * (1.0f / 16777216.0f);
    return 2.0f * u - 1.0f;
}

// Distance with the reference's powf form (PBD_cuda.cu:3-6); used for rest lengths.
float distance_powf(const float* a, const float* b) {
    return powf(powf((a[0] - b[0]), 2) + powf((a[1] - b[1]), 2) + powf((a[2] - b[2]), 2), 0.5f);
}

// PBDObject::initPosition (PBD_cuda.cu:831-852) + initMeshTopolIndices (:854-869),
// then seeded jitter per coordinate and the origin translation.
void make_grid(const pbd_grid_desc& d, std::vector<float>& pos, std::vector<int>& tris) {
    int rx = d.res_x, ry = d.res_y;
    pos.assign((size_t)3 * rx * ry, 0.0f);
    float li = d.size_x / (rx - 1);
    float hi = d.size_y / (ry - 1);
    size_t k = 0;
    for (int i = 0; i < ry; i++)
        for (int j = 0; j < rx; j++) {
            pos[3 * k + 0] = 0.0f + j * li;
            pos[3 * k + 1] = 0.0f;
            pos[3 * k + 2] = 0.0f + i * hi;
            k++;
        }
    if (d.jitter != 0.0f) {
        uint64_t s = d.seed;
        for (size_t i = 0; i < pos.size(); i++) pos[i] = pos[i] + d.jitter * uniform_pm1(&s);
    }
    if (d.origin[0] != 0.0f || d.origin[1] != 0.0f || d.origin[2] != 0.0f)
        for (size_t i = 0; i < pos.size(); i++) pos[i] = pos[i] + d.origin[i % 3];
    tris.clear();
    int num = rx * ry;
    tris.reserve((size_t)6 * (rx - 1) * (ry - 1));
    for (int i = 0; i < num - rx; i++) {
        if (i % rx == rx - 1) continue;
        int q[6] = {i, i + rx, i + rx + 1, i, i + rx + 1, i + 1};
        tris.insert(tris.end(), q, q + 6);
    }
}

// ConstraintPBD::InitDistanceIndices (PBD_cuda.cu:19-54).
void grid_distance_pairs(int rx, int ry, std::vector<int>& pairs) {
    pairs.clear();
    int num = rx * ry;
    auto push = [&](int a, int b) { pairs.push_back(a); pairs.push_back(b); };
    for (int i = 0; i < num - rx; i++) {
        if (i % rx == 0) {
            push(i, i + 1);
            push(i, i + rx);
        } else if (i % rx == rx - 1) {
            push(i, i + rx);
            push(i, i + rx - 1);
        } else {
            push(i, i + 1);
            push(i, i + rx);
            push(i, i + rx - 1);
        }
    }
    for (int i = num - rx; i < num - 1; ++i) push(i, i + 1);
}

static inline uint64_t edge_key(int a, int b) {
    int lo = a < b ? a : b, hi = a < b ? b : a;
    return ((uint64_t)(uint32_t)lo << 32) | (uint32_t)hi;
}

// Distance prims of a general mesh: every triangle edge once, in the order of
// its first sighting (triangles in index order, edges v0v1, v1v2, v2v0).
void mesh_edge_pairs(const std::vector<int>& tris, std::vector<int>& pairs) {
    pairs.clear();
    std::unordered_map<uint64_t, int> seen;
    seen.reserve(tris.size() * 2);
    for (size_t t = 0; t + 2 < tris.size(); t += 3)
        for (int k = 0; k < 3; k++) {
            int a = tris[t + k], b = tris[t + (k + 1) % 3];
            if (seen.emplace(edge_key(a, b), 1).second) {
                pairs.push_back(a);
                pairs.push_back(b);
            }
        }
}

// One dihedral prim per interior edge: (e0, e1, o0, o1) = the edge in the
// first triangle's winding, that triangle's opposite vertex, the second
// triangle's opposite vertex; emitted at the second sighting.  Bridson,
// Marino, Fedkiw 2003, sec. 4.1 (the reference's InitBendingConstr is an
// empty stub, PBD_cuda.cu:110-123).
void build_edge_table(const int* tris, int nt, std::vector<int>& tedge, std::vector<int2>& edges) {
    std::unordered_map<uint64_t, int> id;
    id.reserve((size_t)3 * nt);
    tedge.assign((size_t)3 * nt, -1);
    edges.clear();
    for (int t = 0; t < nt; t++)
        for (int k = 0; k < 3; k++) {
            int a = tris[3 * t + k], b = tris[3 * t + (k + 1) % 3];
            int lo = a < b ? a : b, hi = a < b ? b : a;
            auto it = id.emplace(edge_key(lo, hi), (int)edges.size());
            if (it.second) edges.push_back(make_int2(lo, hi));
            tedge[3 * t + k] = it.first->second;
        }
}

void bending_quads(const std::vector<int>& tris, std::vector<int>& quads) {
    quads.clear();
    struct E { int a, b, o, n; };
    std::unordered_map<uint64_t, E> seen;
    seen.reserve(tris.size() * 2);
    for (size_t t = 0; t + 2 < tris.size(); t += 3)
        for (int k = 0; k < 3; k++) {
            int a = tris[t + k], b = tris[t + (k + 1) % 3], o = tris[t + (k + 2) % 3];
            auto it = seen.find(edge_key(a, b));
            if (it == seen.end()) {
                seen.emplace(edge_key(a, b), E{a, b, o, 1});
            } else if (it->second.n == 1) {
                quads.push_back(it->second.a);
                quads.push_back(it->second.b);
                quads.push_back(it->second.o);
                quads.push_back(o);
                it->second.n = 2;
            }
        }
}

float atan2_poly(float y, float x) { return atan2_hd(y, x); }

float dihedral_angle(const float* e0, const float* e1, const float* o0, const float* o1) {
    return dihedral_hd(mk(e0[0], e0[1], e0[2]), mk(e1[0], e1[1], e1[2]), mk(o0[0], o0[1], o0[2]),
                       mk(o1[0], o1[1], o1[2]));
}

// initConstr (PBD_cuda.cu:798-829): distance prims (primList (2i,2) and rest
// length via the powf Distance, InitDistanceInfo :56-108), then bending prims
// (primList (start,4), rest dihedral angle) when the BENDING bit is set.
// Pins get weight 0 (massBuffer acts as the weight, :1113-1118,1280-1281).
void build_prims(Cloth& c) {
    std::vector<int> pairs;
    if (c.has_user_pairs) pairs = c.user_pairs;
    else if (c.rx > 0) grid_distance_pairs(c.rx, c.ry, pairs);
    else mesh_edge_pairs(c.tris, pairs);
    if (!(c.ct & PBD_DISTANCE)) pairs.clear();
    std::vector<int> quads;
    if (c.ct & PBD_BENDING) {
        if (c.has_user_quads) quads = c.user_quads;
        else bending_quads(c.tris, quads);
    }
    c.ne = (int)pairs.size() / 2;
    c.nb = (int)quads.size() / 4;
    c.prim_list.clear();
    c.indices.clear();
    c.rest.clear();
    c.prim_list.reserve((size_t)2 * c.np());
    c.indices.reserve(pairs.size() + quads.size());
    c.rest.reserve(c.np());
    for (int i = 0; i < c.ne; i++) {
        int a = pairs[2 * i], b = pairs[2 * i + 1];
        PBD_CHECK(a >= 0 && a < c.nv && b >= 0 && b < c.nv && a != b, PBD_ERR_ARG,
                  "distance prim vertex out of range");
        c.prim_list.push_back((int)c.indices.size());
        c.prim_list.push_back(2);
        c.indices.push_back(a);
        c.indices.push_back(b);
        c.rest.push_back(distance_powf(&c.pos0[3 * a], &c.pos0[3 * b]));
    }
    for (int i = 0; i < c.nb; i++) {
        const int* q = &quads[4 * i];
        for (int k = 0; k < 4; k++) {
            PBD_CHECK(q[k] >= 0 && q[k] < c.nv, PBD_ERR_ARG, "bending prim vertex out of range");
            for (int j = 0; j < k; j++) PBD_CHECK(q[j] != q[k], PBD_ERR_ARG, "bending prim repeats a vertex");
        }
        c.prim_list.push_back((int)c.indices.size());
        c.prim_list.push_back(4);
        for (int k = 0; k < 4; k++) c.indices.push_back(q[k]);
        c.rest.push_back(dihedral_angle(&c.pos0[3 * q[0]], &c.pos0[3 * q[1]], &c.pos0[3 * q[2]],
                                        &c.pos0[3 * q[3]]));
    }
    if (!c.user_rest.empty()) {
        PBD_CHECK((int)c.user_rest.size() == c.np(), PBD_ERR_ARG,
                  "explicit rest values do not match the cloth's constraint set");
        c.rest = c.user_rest;
    }
    if (c.w_in.empty()) {
        c.w.assign(c.nv, 1.0f);
    } else {
        PBD_CHECK((int)c.w_in.size() == c.nv, PBD_ERR_ARG, "explicit weights do not match the vertex count");
        c.w = c.w_in;
    }
    for (int v : c.pins) {
        PBD_CHECK(v >= 0 && v < c.nv, PBD_ERR_ARG, "pin index out of range");
        c.w[v] = 0.0f;
    }
}

// GenePoint2PrimsMap + GenePrim2PrimsMap (PBD_cuda.cu:155-202): for prim p,
// for each of its points in order, every other prim on that point in
// ascending prim id; duplicates kept.  Built as a CSR in O(sum deg^2).
void build_adjacency(Cloth& c) {
    int np = c.np();
    std::vector<int> cnt((size_t)c.nv + 1, 0);
    for (int p = 0; p < np; p++)
        for (int k = 0; k < c.prim_list[2 * p + 1]; k++) cnt[c.indices[c.prim_list[2 * p] + k] + 1]++;
    for (int v = 0; v < c.nv; v++) cnt[v + 1] += cnt[v];
    std::vector<int> fill(cnt.begin(), cnt.end() - 1);
    std::vector<int> p2p((size_t)cnt[c.nv]);
    for (int p = 0; p < np; p++)
        for (int k = 0; k < c.prim_list[2 * p + 1]; k++) p2p[fill[c.indices[c.prim_list[2 * p] + k]]++] = p;
    c.adj_start_num.assign((size_t)2 * np, 0);
    c.adj.clear();
    for (int p = 0; p < np; p++) {
        int start = (int)c.adj.size();
        for (int k = 0; k < c.prim_list[2 * p + 1]; k++) {
            int v = c.indices[c.prim_list[2 * p] + k];
            for (int j = cnt[v]; j < cnt[v + 1]; j++)
                if (p2p[j] != p) c.adj.push_back(p2p[j]);
        }
        c.adj_start_num[2 * p] = start;
        c.adj_start_num[2 * p + 1] = (int)c.adj.size() - start;
    }
}

// UT::splitString semantics (Utility.h:6-28): split on `sep`, keep a trailing
// piece only when non-empty.
static std::vector<std::string> split(const std::string& s, char sep) {
    std::vector<std::string> out;
    size_t a = 0, b;
    while ((b = s.find(sep, a)) != std::string::npos) {
        out.push_back(s.substr(a, b - a));
        a = b + 1;
    }
    if (a != s.size()) out.push_back(s.substr(a));
    return out;
}

// IO::ReadTopolFromTxt (PBD_Basic.cuh:532-601): line 0 "label:x,y,z,x,y,z,...",
// line 1 "label:i,j,k,..." (triangles).  IO::ReadDistConstrFromTxt (:605-688):
// line 0 skipped, every later line "label:i0,i1,i0,i1,...".
void read_txt_cloth(const char* topol, const char* constr, std::vector<float>& pos,
                    std::vector<int>& tris, std::vector<int>& pairs, bool& has_pairs) {
    std::ifstream in(topol);
    PBD_CHECK(in.is_open(), PBD_ERR_IO, std::string("cannot open ") + topol);
    pos.clear();
    tris.clear();
    std::string line;
    int ln = 0;
    try {
        while (std::getline(in, line)) {
            if (!line.empty() && line.back() == '\r') line.pop_back();
            auto parts = split(line, ':');
            PBD_CHECK(parts.size() >= 2, PBD_ERR_IO, "topology line without ':'");
            auto vals = split(parts[1], ',');
            if (ln == 0) {
                PBD_CHECK(vals.size() % 3 == 0, PBD_ERR_IO, "position count not a multiple of 3");
                for (auto& v : vals) pos.push_back(std::stof(v));
            } else {
                PBD_CHECK(vals.size() % 3 == 0, PBD_ERR_IO, "index count not a multiple of 3");
                for (auto& v : vals) tris.push_back(std::stoi(v));
            }
            ++ln;
        }
    } catch (const std::logic_error&) {
        throw Error(PBD_ERR_IO, std::string("malformed number in ") + topol);
    }
    has_pairs = false;
    pairs.clear();
    if (constr && constr[0]) {
        std::ifstream ic(constr);
        PBD_CHECK(ic.is_open(), PBD_ERR_IO, std::string("cannot open ") + constr);
        ln = 0;
        try {
            while (std::getline(ic, line)) {
                if (ln++ == 0) continue;
                if (!line.empty() && line.back() == '\r') line.pop_back();
                auto parts = split(line, ':');
                PBD_CHECK(parts.size() >= 2, PBD_ERR_IO, "constraint line without ':'");
                auto vals = split(parts[1], ',');
                PBD_CHECK(vals.size() % 2 == 0, PBD_ERR_IO, "constraint index count odd");
                for (auto& v : vals) pairs.push_back(std::stoi(v));
            }
        } catch (const std::logic_error&) {
            throw Error(PBD_ERR_IO, std::string("malformed number in ") + constr);
        }
        has_pairs = true;
    }
}

// ---------------------------------------------------------------------------
// Tile planner.
//
// A tile is a set of interior vertices I.  Running P = sweeps*C colour phases
// on the tile alone reproduces the global colour order exactly on I when the
// loaded set L contains the backward dependency cone of I: going backwards
// over phases k = P-1..0 with colour c_k, D_k = D_{k+1} plus every vertex of a
// colour-c_k prim incident to D_{k+1} (each vertex has at most one prim per
// colour).  Only the colour-c_k prims incident to D_{k+1} need processing at
// phase k; others only disturb vertices outside D_{k+1}, which no later phase
// reads.  With entry(v) = the phase at which v joins the cone (P for I), a
// prim of colour c is needed at phase s*C + c iff some vertex has entry
// > s*C + c, i.e. for sweeps s <= S(p) = floor((max entry - 1 - c) / C): the
// needed prims of a colour at sweep s are those with S(p) >= s.  Each colour's
// prims are stored in decreasing S(p), so phase (s, c) runs a prefix of the
// colour's list; prims of one colour share no vertex, so their order within
// the colour changes nothing.  A launch of n < sweeps sweeps has the cone of
// the last n sweeps: its sweep s uses the prefix of sweep s + sweeps - n.
// The owned distance prims (first vertex in I) are added to D_P so that the
// residual can be evaluated on valid values at the end of the launch.
// L is listed in the order vertices join the cone (interior and owned first,
// then by decreasing entry).  A prim needed at phase k joins its vertices to
// the cone at phase k at the latest, so both have entry >= k: a launch of
// the last n sweeps reads only vertices with entry >= (sweeps - n) * C, a
// prefix of L whose length the tile records.
// ---------------------------------------------------------------------------
namespace {

struct ClothTopo {
    const Cloth* c;
    std::vector<int> vp_start, vp;   // vertex -> prims
    std::vector<int> vcp;            // vertex x colour -> prim or -1
    int C;
};

void build_topo(const Cloth& c, ClothTopo& t) {
    t.c = &c;
    t.C = c.ncolors;
    int np = c.np();
    t.vp_start.assign((size_t)c.nv + 1, 0);
    for (int p = 0; p < np; p++)
        for (int k = 0; k < c.prim_list[2 * p + 1]; k++) t.vp_start[c.indices[c.prim_list[2 * p] + k] + 1]++;
    for (int v = 0; v < c.nv; v++) t.vp_start[v + 1] += t.vp_start[v];
    std::vector<int> fill(t.vp_start.begin(), t.vp_start.end() - 1);
    t.vp.resize((size_t)t.vp_start[c.nv]);
    for (int p = 0; p < np; p++)
        for (int k = 0; k < c.prim_list[2 * p + 1]; k++) t.vp[fill[c.indices[c.prim_list[2 * p] + k]]++] = p;
    t.vcp.assign((size_t)c.nv * t.C, -1);
    for (int v = 0; v < c.nv; v++)
        for (int j = t.vp_start[v]; j < t.vp_start[v + 1]; j++) t.vcp[(size_t)v * t.C + c.color[t.vp[j]]] = t.vp[j];
}

struct TileOut {
    std::vector<int> load;   // local -> cloth vertex (interior first)
    int n_int = 0;
    std::vector<std::vector<int>> dist_by_col, bend_by_col;  // cloth prim ids, by decreasing last needed sweep
    std::vector<int> dcount, bcount;  // [s * C + c]: prims of colour c needed at sweep s (a prefix); empty: all
    std::vector<int> lcount;  // [n - 1]: loaded vertices a launch of the last n sweeps needs (a prefix of load)
    std::vector<int> own;    // owned distance prims
};

// Computes the loaded set of one tile.  Returns its size; fills `out` only
// when it fits in max_load.
int plan_one(const ClothTopo& t, const std::vector<int>& interior, int phases, int max_load,
             std::vector<int>& mark, std::vector<int>& entry, int stamp, TileOut* out) {
    const Cloth& c = *t.c;
    std::vector<int> D(interior);
    D.reserve((size_t)max_load + 64);
    for (int v : interior) {
        mark[v] = stamp;
        entry[v] = phases;
    }
    std::vector<int> own;
    {
        std::vector<char> is_int;  // membership via mark (stamp) of interior only
        for (int v : interior)
            for (int j = t.vp_start[v]; j < t.vp_start[v + 1]; j++) {
                int p = t.vp[j];
                if (c.prim_list[2 * p + 1] != 2) continue;
                if (c.indices[c.prim_list[2 * p]] != v) continue;
                own.push_back(p);
            }
        for (int p : own)
            for (int k = 0; k < 2; k++) {
                int u = c.indices[c.prim_list[2 * p] + k];
                if (mark[u] != stamp) { mark[u] = stamp; entry[u] = phases; D.push_back(u); }
            }
    }
    std::vector<int> frontier(D), nf;
    for (int k = phases - 1; k >= 0 && !frontier.empty(); k--) {
        int col = k % t.C;
        nf.clear();
        for (int v : frontier) {
            int p = t.vcp[(size_t)v * t.C + col];
            if (p >= 0) {
                int st = c.prim_list[2 * p], cnt = c.prim_list[2 * p + 1];
                for (int q = 0; q < cnt; q++) {
                    int u = c.indices[st + q];
                    if (mark[u] != stamp) { mark[u] = stamp; entry[u] = k; D.push_back(u); nf.push_back(u); }
                }
            }
            bool outside = false;
            for (int j = t.vp_start[v]; j < t.vp_start[v + 1] && !outside; j++) {
                int pp = t.vp[j];
                int st = c.prim_list[2 * pp], cnt = c.prim_list[2 * pp + 1];
                for (int q = 0; q < cnt; q++)
                    if (mark[c.indices[st + q]] != stamp) { outside = true; break; }
            }
            if (outside) nf.push_back(v);
        }
        frontier.swap(nf);
        if ((int)D.size() > max_load) return (int)D.size();
    }
    if ((int)D.size() > max_load || !out) return (int)D.size();
    out->load = D;
    out->n_int = (int)interior.size();
    out->own = own;
    out->dist_by_col.assign(t.C, {});
    out->bend_by_col.assign(t.C, {});
    const int sweeps = t.C > 0 ? phases / t.C : 0;
    // (last needed sweep, prim) per colour; prims needed at no phase are dropped
    std::vector<std::vector<std::pair<int, int>>> dl(t.C), bl(t.C);
    for (int v : D)
        for (int j = t.vp_start[v]; j < t.vp_start[v + 1]; j++) {
            int p = t.vp[j];
            int st = c.prim_list[2 * p], cnt = c.prim_list[2 * p + 1];
            if (c.indices[st] != v) continue;  // visit each prim once, from its first vertex
            bool all = true;
            int emax = -1;
            for (int q = 0; q < cnt; q++) {
                const int u = c.indices[st + q];
                if (mark[u] != stamp) { all = false; break; }
                emax = std::max(emax, entry[u]);
            }
            if (!all) continue;
            const int col = c.color[p];
            const int last = emax - 1 - col;
            if (last < 0) continue;
            (cnt == 2 ? dl : bl)[col].push_back({last / t.C, p});
        }
    auto order = [&](std::vector<std::vector<std::pair<int, int>>>& lists, std::vector<std::vector<int>>& by_col,
                     std::vector<int>& count) {
        count.assign((size_t)sweeps * t.C, 0);
        for (int col = 0; col < t.C; col++) {
            auto& l = lists[col];
            std::sort(l.begin(), l.end(), [](const std::pair<int, int>& a, const std::pair<int, int>& b) {
                return a.first != b.first ? a.first > b.first : a.second < b.second;
            });
            for (auto& e : l) {
                by_col[col].push_back(e.second);
                for (int s = 0; s <= e.first && s < sweeps; s++) count[(size_t)s * t.C + col]++;
            }
        }
    };
    order(dl, out->dist_by_col, out->dcount);
    order(bl, out->bend_by_col, out->bcount);
    for (size_t m = 1; m < D.size(); m++)
        PBD_CHECK(entry[D[m]] <= entry[D[m - 1]], PBD_ERR_STATE, "internal: cone list not in entry order");
    out->lcount.assign((size_t)sweeps, (int)D.size());
    for (int n = 1; n < sweeps; n++) {
        const int from = (sweeps - n) * t.C;
        int m = 0;
        while (m < (int)D.size() && entry[D[m]] >= from) m++;
        out->lcount[(size_t)n - 1] = m;
    }
    return (int)D.size();
}

// Interior sets: rectangular blocks for grid cloths, contiguous id ranges
// otherwise; any block whose cone does not fit is split in two.
struct Block { int x0, x1, y0, y1; };

}  // namespace

void plan_tiles(std::vector<Cloth>& cloths, int max_load, int sweeps, bool split_whole, TilePlan& plan) {
    plan = TilePlan();
    struct Pending { int cloth; std::vector<TileOut> outs; };
    std::vector<Pending> per(cloths.size());
    std::vector<std::thread> pool;
    unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    // Large cloths are planned in parallel; a small one is a single tile of
    // all its vertices (no halo), whole unless split_whole asks for it to be
    // stepped like a multi-tile cloth.
    std::vector<int> work;
    for (size_t ci = 0; ci < cloths.size(); ci++) {
        Cloth& c = cloths[ci];
        per[ci].cloth = (int)ci;
        const bool fits = c.nv <= max_load;
        c.whole = fits && !split_whole;
        c.max_loaded = 0;
        if (fits) {
            TileOut o;
            o.load.resize(c.nv);
            for (int v = 0; v < c.nv; v++) o.load[v] = v;
            o.n_int = c.nv;
            o.dist_by_col.assign(c.ncolors, {});
            o.bend_by_col.assign(c.ncolors, {});
            for (int j = 0; j < c.np(); j++) {
                int p = c.sorted[j];
                (c.prim_list[2 * p + 1] == 2 ? o.dist_by_col : o.bend_by_col)[c.color[p]].push_back(p);
                if (c.prim_list[2 * p + 1] == 2) o.own.push_back(p);
            }
            std::sort(o.own.begin(), o.own.end());
            per[ci].outs.push_back(std::move(o));
            c.max_loaded = c.nv;
        } else {
            work.push_back((int)ci);
        }
    }
    std::string err;
    auto run_cloth = [&](int ci) {
        Cloth& c = cloths[ci];
        ClothTopo topo;
        build_topo(c, topo);
        std::vector<int> mark((size_t)c.nv, -1), entry((size_t)c.nv, 0);
        int stamp = 0;
        int phases = std::max(1, sweeps) * c.ncolors;
        std::vector<Block> todo;
        if (c.rx > 0) {
            int T = std::max(4, (int)std::sqrt((double)max_load) - 2);
            // shrink the nominal block until a central block's cone fits
            for (;;) {
                int cx = std::max(0, c.rx / 2 - T / 2), cy = std::max(0, c.ry / 2 - T / 2);
                std::vector<int> in;
                for (int y = cy; y < std::min(c.ry, cy + T); y++)
                    for (int x = cx; x < std::min(c.rx, cx + T); x++) in.push_back(y * c.rx + x);
                int n = plan_one(topo, in, phases, max_load, mark, entry, stamp++, nullptr);
                if (n <= max_load || T <= 4) break;
                T = std::max(4, (int)(T * 0.85));
            }
            for (int y0 = 0; y0 < c.ry; y0 += T)
                for (int x0 = 0; x0 < c.rx; x0 += T)
                    todo.push_back({x0, std::min(c.rx, x0 + T), y0, std::min(c.ry, y0 + T)});
        } else {
            int T = std::max(16, max_load / 3);
            for (int a = 0; a < c.nv; a += T) todo.push_back({a, std::min(c.nv, a + T), 0, 1});
        }
        std::vector<TileOut> outs;
        while (!todo.empty()) {
            Block b = todo.back();
            todo.pop_back();
            std::vector<int> in;
            if (c.rx > 0) {
                for (int y = b.y0; y < b.y1; y++)
                    for (int x = b.x0; x < b.x1; x++) in.push_back(y * c.rx + x);
            } else {
                for (int v = b.x0; v < b.x1; v++) in.push_back(v);
            }
            TileOut o;
            int n = plan_one(topo, in, phases, max_load, mark, entry, stamp++, &o);
            if (n <= max_load) {
                outs.push_back(std::move(o));
                c.max_loaded = std::max(c.max_loaded, n);
                continue;
            }
            int wx = b.x1 - b.x0, wy = b.y1 - b.y0;
            if ((c.rx > 0 && wx * wy <= 1) || (c.rx == 0 && wx <= 1)) {
                err = "a single vertex's dependency cone exceeds the tile budget; lower the sweeps per launch";
                return;
            }
            if (c.rx > 0 && wx >= wy) {
                int m = b.x0 + wx / 2;
                todo.push_back({b.x0, m, b.y0, b.y1});
                todo.push_back({m, b.x1, b.y0, b.y1});
            } else if (c.rx > 0) {
                int m = b.y0 + wy / 2;
                todo.push_back({b.x0, b.x1, b.y0, m});
                todo.push_back({b.x0, b.x1, m, b.y1});
            } else {
                int m = b.x0 + wx / 2;
                todo.push_back({b.x0, m, 0, 1});
                todo.push_back({m, b.x1, 0, 1});
            }
        }
        // spatial order: by the first interior vertex
        std::sort(outs.begin(), outs.end(), [](const TileOut& a, const TileOut& b) { return a.load[0] < b.load[0]; });
        per[ci].outs = std::move(outs);
    };
    std::atomic_size_t next{0};
    for (unsigned t = 0; t < hw; t++)
        pool.emplace_back([&]() {
            for (;;) {
                size_t i = next.fetch_add(1);
                if (i >= work.size()) break;
                run_cloth(work[i]);
            }
        });
    for (auto& th : pool) th.join();
    PBD_CHECK(err.empty(), PBD_ERR_LIMIT, err);

    // Flatten.
    for (size_t ci = 0; ci < cloths.size(); ci++) {
        Cloth& c = cloths[ci];
        c.tile_begin = (int)plan.tiles.size();
        c.tiles_loaded = c.tiles_interior = c.tiles_owned = 0;
        c.tile_record_bytes.assign((size_t)std::max(1, sweeps) + 1, 0.0);
        c.tile_loaded_by_sweeps.assign((size_t)std::max(1, sweeps) + 1, 0.0);
        for (auto& o : per[ci].outs) {
            TileDev td{};
            td.cloth = (int)ci;
            td.whole = c.whole ? 1 : 0;
            td.load_off = (int)plan.verts.size();
            td.n_load = (int)o.load.size();
            td.n_int = o.n_int;
            td.ncolors = c.ncolors;
            td.col_off = (int)plan.cols.size();
            td.dist_off = (int)plan.dist_rest.size();
            td.bend_off = (int)plan.bend_rest.size();
            td.own_off = (int)plan.own_rest.size();
            td.n_own = (int)o.own.size();
            std::vector<int> local((size_t)c.nv, -1);
            for (size_t k = 0; k < o.load.size(); k++) {
                local[o.load[k]] = (int)k;
                plan.verts.push_back(c.vbase + o.load[k]);
            }
            PBD_CHECK(o.load.size() <= 65535, PBD_ERR_LIMIT, "tile exceeds 16-bit local indices");
            int dsum = 0, bsum = 0;
            for (int col = 0; col < c.ncolors; col++) {
                plan.cols.push_back(dsum);
                dsum += (int)o.dist_by_col[col].size();
            }
            plan.cols.push_back(dsum);
            for (int col = 0; col < c.ncolors; col++) {
                plan.cols.push_back(bsum);
                bsum += (int)o.bend_by_col[col].size();
            }
            plan.cols.push_back(bsum);
            // per-sweep prefixes of the partial tiles: sweeps x C distance
            // counts, sweeps x C bending counts, then the loaded vertices of
            // a launch of 1 .. sweeps sweeps
            td.plan_sweeps = o.dcount.empty() ? 0 : (int)(o.dcount.size() / std::max(1, c.ncolors));
            plan.cols.insert(plan.cols.end(), o.dcount.begin(), o.dcount.end());
            plan.cols.insert(plan.cols.end(), o.bcount.begin(), o.bcount.end());
            if (td.plan_sweeps) {
                PBD_CHECK((int)o.lcount.size() == td.plan_sweeps, PBD_ERR_STATE, "internal: load prefix count");
                plan.cols.insert(plan.cols.end(), o.lcount.begin(), o.lcount.end());
                for (int n = 1; n <= td.plan_sweeps && n < (int)c.tile_loaded_by_sweeps.size(); n++)
                    c.tile_loaded_by_sweeps[n] += o.lcount[(size_t)n - 1];
            } else {
                for (size_t n = 1; n < c.tile_loaded_by_sweeps.size(); n++)
                    c.tile_loaded_by_sweeps[n] += (double)o.load.size();
            }
            c.tiles_loaded += (long long)o.load.size();
            c.tiles_interior += o.n_int;
            c.tiles_owned += (long long)o.own.size();
            if (o.dcount.empty()) {
                // every prim at every sweep: the bytes of one sweep
                double b = 0.0;
                for (int col = 0; col < c.ncolors; col++)
                    b += 8.0 * o.dist_by_col[col].size() + 12.0 * o.bend_by_col[col].size();
                c.tile_record_bytes[1] += b;
            } else {
                const int ps = td.plan_sweeps;
                for (int n = 1; n <= ps && n < (int)c.tile_record_bytes.size(); n++) {
                    double b = 0.0;
                    for (int sw = 0; sw < n; sw++)
                        for (int col = 0; col < c.ncolors; col++) {
                            const size_t q = (size_t)(sw + ps - n) * c.ncolors + col;
                            b += 8.0 * o.dcount[q] + 12.0 * o.bcount[q];
                        }
                    c.tile_record_bytes[n] += b;
                }
            }
            plan.stat_loaded += (long long)o.load.size();
            plan.stat_interior += o.n_int;
            for (int col = 0; col < c.ncolors; col++) plan.stat_dist_loaded += (long long)o.dist_by_col[col].size();
            for (size_t q = 0; q < o.dcount.size(); q++) plan.stat_dist_phase_needed += o.dcount[q];
            if (o.dcount.empty())
                for (int col = 0; col < c.ncolors; col++)
                    plan.stat_dist_phase_needed += (long long)sweeps * o.dist_by_col[col].size();
            for (int col = 0; col < c.ncolors; col++)
                for (int p : o.dist_by_col[col]) {
                    const int* ix = &c.indices[c.prim_list[2 * p]];
                    plan.dist.push_back((uint32_t)local[ix[0]] | ((uint32_t)local[ix[1]] << 16));
                    plan.dist_rest.push_back(c.rest[p]);
                }
            for (int col = 0; col < c.ncolors; col++)
                for (int p : o.bend_by_col[col]) {
                    const int* ix = &c.indices[c.prim_list[2 * p]];
                    plan.bend.push_back((uint32_t)local[ix[0]] | ((uint32_t)local[ix[1]] << 16));
                    plan.bend.push_back((uint32_t)local[ix[2]] | ((uint32_t)local[ix[3]] << 16));
                    plan.bend_rest.push_back(c.rest[p]);
                }
            for (int p : o.own) {
                const int* ix = &c.indices[c.prim_list[2 * p]];
                plan.own.push_back((uint32_t)local[ix[0]] | ((uint32_t)local[ix[1]] << 16));
                plan.own_rest.push_back(c.rest[p]);
            }
            plan.tiles.push_back(td);
        }
        c.tile_end = (int)plan.tiles.size();
    }
    // Tile neighbours: t and n are neighbours when one loads a vertex of the
    // other's interior (a later chunk of t reads what n wrote, or writes what
    // n read).  Only tiles of the same cloth share vertices.
    {
        const size_t nt = plan.tiles.size();
        std::vector<std::vector<int>> nb(nt);
        int nv_all = 0;
        for (const Cloth& c : cloths) nv_all = std::max(nv_all, c.vbase + c.nv);
        std::vector<int> owner((size_t)nv_all, -1);
        for (size_t t = 0; t < nt; t++) {
            const TileDev& td = plan.tiles[t];
            if (td.whole) continue;
            for (int k = 0; k < td.n_int; k++) owner[plan.verts[(size_t)td.load_off + k]] = (int)t;
        }
        for (size_t t = 0; t < nt; t++) {
            const TileDev& td = plan.tiles[t];
            if (td.whole) continue;
            nb[t].push_back((int)t);
            for (int k = td.n_int; k < td.n_load; k++) {
                const int n = owner[plan.verts[(size_t)td.load_off + k]];
                if (n >= 0 && n != (int)t) {
                    nb[t].push_back(n);
                    nb[n].push_back((int)t);
                }
            }
        }
        plan.nbr_off.assign(nt + 1, 0);
        plan.nbr.clear();
        for (size_t t = 0; t < nt; t++) {
            std::sort(nb[t].begin(), nb[t].end());
            nb[t].erase(std::unique(nb[t].begin(), nb[t].end()), nb[t].end());
            plan.nbr.insert(plan.nbr.end(), nb[t].begin(), nb[t].end());
            plan.nbr_off[t + 1] = (int)plan.nbr.size();
        }
    }
    // XCD-aware launch order: consecutive (spatially adjacent) tiles go to the
    // same XCD.  Workgroups are dealt to the 8 XCDs round-robin by block id, so
    // the j-th tile of group g gets block id j*8 + g.
    auto xcd_order = [](const std::vector<int>& ids) {
        const int X = 8;
        int n = (int)ids.size();
        int per_g = (n + X - 1) / X;
        std::vector<int> ord((size_t)per_g * X, -1);
        for (int i = 0; i < n; i++) {
            int g = i / per_g, j = i % per_g;
            ord[(size_t)j * X + g] = ids[i];
        }
        return ord;
    };
    std::vector<int> whole_ids, part_ids;
    for (int i = 0; i < (int)plan.tiles.size(); i++) (plan.tiles[i].whole ? whole_ids : part_ids).push_back(i);
    plan.whole_order = xcd_order(whole_ids);
    plan.partial_order = xcd_order(part_ids);
}

}  // namespace pbd
