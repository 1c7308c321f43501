// The reference's text cache format: one buffer per line,
//
//     <name>|<type name>|<item>;<item>;...;
//
// with a float3 item written "x,y,z", an int2 item "a,b" and scalars as the
// C++ stream writes them (Buffer<T>::Save, Buffer.cuh:70-80; the float3/int2
// inserters, PBD_Basic.cuh:68-78; IO::SaveData, PBD_Basic.cuh:278-296, which
// writes one item without the trailing ';').  Two files describe a cloth:
//
//   mesh file (PBDObject::SaveMeshTopol, PBD_cuda.cu:907-925): Indices (the
//     triangles), P (current positions), primList (start, 3 per triangle),
//     velBuffer, massBuffer, dampingRate, gravity;
//   constraint file (PBDObject::SaveConstraint, PBD_cuda.cu:927-938, through
//     ConstraintPBD::Save, PBD_cuda.cu:627-640): Indices and primList of the
//     constraint prims, P (the positions the constraints were built from),
//     prdPBuffer (predicted positions), restPosBuffer (anchor positions),
//     restLengthBuffer (one per distance prim), stiffnessBuffer and
//     constraintType (one per prim: DISTANCE 1, BENDING 16, ANCHOR 256).
//
// The reference builds no bending prims (ConstraintPBD::InitBendingConstr,
// PBD_cuda.cu:107-120, is empty), so it has no line for their rest angles;
// this library writes them as "restAngleBuffer" (the name of the reference's
// unused ConstraintPBD::restAngleBuffer, PBD_Basic.cuh:103), a line the
// reference's reader skips as it skips every unknown name.
//
// Reading follows IO::ReadTopolFromCache / IO::ReadConstraintFromCache
// (PBD_Basic.cuh:326-530): a line is split on '|' and used only when it has
// exactly three pieces, the items are split on ';' and the components on ','
// with UT::splitString (Utility.h:8-28), which drops an empty last piece; so
// an empty buffer ("name|type|") is skipped like a missing one.  Unlike the
// reference, a malformed number or an inconsistent count is an error here
// (the reference's std::stof/std::stoi parse a prefix or throw), and the
// restPosBuffer items are not appended to the predicted positions (the
// reference's reader does that, PBD_Basic.cuh:466-473; its writer's
// restPosBuffer holds the anchors' positions, not predicted positions).
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <string>
#include <vector>

#include "pbd_internal.hpp"

namespace pbd {

namespace {

// UT::splitString (Utility.h:8-28): pieces between separators; a last piece
// is kept only when non-empty.
std::vector<std::string> split_pieces(const std::string& s, char sep) {
    std::vector<std::string> out;
    size_t a = 0, b;
    while ((b = s.find(sep, a)) != std::string::npos) {
        out.push_back(s.substr(a, b - a));
        a = b + 1;
    }
    if (a != s.size()) out.push_back(s.substr(a));
    return out;
}

// --- writing ---------------------------------------------------------------

struct Writer {
    std::ofstream os;
    int digits;
    char buf[64];
    Writer(const char* path, int precision) : os(path, std::ios::out | std::ios::trunc), digits(precision) {
        PBD_CHECK(os.is_open(), PBD_ERR_IO, std::string("cannot write ") + path);
    }
    // operator<<(std::ostream&, float) with the stream's default flags: %g
    // at `digits` significant digits of the value widened to double
    const char* f(float v) {
        std::snprintf(buf, sizeof buf, "%.*g", digits, (double)v);
        return buf;
    }
    void head(const char* name, const char* type) { os << name << '|' << type << '|'; }
    void floats(const char* name, const float* v, size_t n) {
        head(name, "float");
        for (size_t i = 0; i < n; i++) os << f(v[i]) << ';';
        os << '\n';
    }
    void ints(const char* name, const int* v, size_t n) {
        head(name, "int");
        for (size_t i = 0; i < n; i++) os << v[i] << ';';
        os << '\n';
    }
    void int2s(const char* name, const int* v, size_t n) {
        head(name, "struct int2");
        for (size_t i = 0; i < n; i++) os << v[2 * i] << ',' << v[2 * i + 1] << ';';
        os << '\n';
    }
    void float3_item(const float* v) {
        os << f(v[0]) << ',';
        os << f(v[1]) << ',';
        os << f(v[2]);
    }
    void float3s(const char* name, const float* v, size_t n) {
        head(name, "struct float3");
        for (size_t i = 0; i < n; i++) {
            float3_item(v + 3 * i);
            os << ';';
        }
        os << '\n';
    }
    void finish(const char* path) {
        os.flush();
        PBD_CHECK(os.good(), PBD_ERR_IO, std::string("write failed: ") + path);
    }
};

// --- reading ---------------------------------------------------------------

using Lines = std::map<std::string, std::vector<std::string>>;  // name -> items

Lines read_lines(const char* path) {
    std::ifstream in(path);
    PBD_CHECK(in.is_open(), PBD_ERR_IO, std::string("cannot open ") + path);
    Lines out;
    std::string line;
    while (std::getline(in, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        auto parts = split_pieces(line, '|');
        if (parts.size() != 3) continue;
        // a name that appears again appends, as the reference's push_back does
        auto items = split_pieces(parts[2], ';');
        auto& dst = out[parts[0]];
        dst.insert(dst.end(), items.begin(), items.end());
    }
    return out;
}

float parse_float(const std::string& s, const char* what) {
    const char* b = s.c_str();
    char* e = nullptr;
    float v = std::strtof(b, &e);  // an underflow returns the subnormal or zero, kept
    PBD_CHECK(e != b && *e == '\0', PBD_ERR_IO, std::string("malformed number in ") + what);
    PBD_CHECK(std::isfinite(v), PBD_ERR_IO, std::string("non-finite number in ") + what);
    return v;
}

int parse_int(const std::string& s, const char* what) {
    const char* b = s.c_str();
    char* e = nullptr;
    errno = 0;
    long v = std::strtol(b, &e, 10);
    PBD_CHECK(e != b && *e == '\0' && errno != ERANGE && v >= -(1l << 31) && v < (1l << 31), PBD_ERR_IO,
              std::string("malformed integer in ") + what);
    return (int)v;
}

bool has(const Lines& L, const char* name) { return L.count(name) != 0; }

std::vector<float> get_floats(const Lines& L, const char* name) {
    std::vector<float> out;
    auto it = L.find(name);
    if (it == L.end()) return out;
    for (auto& s : it->second) out.push_back(parse_float(s, name));
    return out;
}

std::vector<int> get_ints(const Lines& L, const char* name) {
    std::vector<int> out;
    auto it = L.find(name);
    if (it == L.end()) return out;
    for (auto& s : it->second) out.push_back(parse_int(s, name));
    return out;
}

// k components per item, split on ','
std::vector<float> get_vecf(const Lines& L, const char* name, int k) {
    std::vector<float> out;
    auto it = L.find(name);
    if (it == L.end()) return out;
    for (auto& s : it->second) {
        auto c = split_pieces(s, ',');
        PBD_CHECK((int)c.size() == k, PBD_ERR_IO, std::string("wrong component count in ") + name);
        for (auto& x : c) out.push_back(parse_float(x, name));
    }
    return out;
}

std::vector<int> get_veci(const Lines& L, const char* name, int k) {
    std::vector<int> out;
    auto it = L.find(name);
    if (it == L.end()) return out;
    for (auto& s : it->second) {
        auto c = split_pieces(s, ',');
        PBD_CHECK((int)c.size() == k, PBD_ERR_IO, std::string("wrong component count in ") + name);
        for (auto& x : c) out.push_back(parse_int(x, name));
    }
    return out;
}

}  // namespace

void write_cache(const Cloth& c, const float* x, const float* v, const float* p, const char* mesh_path,
                 const char* constr_path, int precision) {
    int digits = precision > 0 ? precision : 6;
    PBD_CHECK(digits <= 17, PBD_ERR_ARG, "precision must be at most 17 digits");
    if (mesh_path && mesh_path[0]) {
        Writer w(mesh_path, digits);
        w.ints("Indices", c.tris.data(), c.tris.size());
        w.float3s("P", x, c.nv);
        std::vector<int> prims((size_t)2 * c.nt);
        for (int t = 0; t < c.nt; t++) {
            prims[2 * t] = 3 * t;
            prims[2 * t + 1] = 3;
        }
        w.int2s("primList", prims.data(), c.nt);
        w.float3s("velBuffer", v, c.nv);
        // massBuffer is the weight the reference's projection divides by
        // (PBD_cuda.cu:1113-1118): the inverse mass, 0 for a pinned vertex
        w.floats("massBuffer", c.w.data(), c.w.size());
        w.head("dampingRate", "float");
        w.os << w.f(c.prm.damping) << '\n';
        w.head("gravity", "struct float3");
        w.float3_item(c.prm.gravity);
        w.os << '\n';
        w.finish(mesh_path);
    }
    if (constr_path && constr_path[0]) {
        // prims in library order (distance, then bending), then one anchor per
        // pinned vertex (weight 0), in vertex order, as ConstraintPBD::
        // InitAnchorConstr adds them (PBD_cuda.cu:122-146: one index, its
        // position in restPosBuffer, stiffness -1)
        std::vector<int> prims(c.prim_list), idx(c.indices), type;
        std::vector<float> stiff, rest_len(c.rest.begin(), c.rest.begin() + c.ne),
            rest_ang(c.rest.begin() + c.ne, c.rest.end()), rest_pos;
        for (int i = 0; i < c.ne; i++) {
            type.push_back(PBD_DISTANCE);
            stiff.push_back(c.stiff[0]);
        }
        for (int i = 0; i < c.nb; i++) {
            type.push_back(PBD_BENDING);
            stiff.push_back(c.stiff[1]);
        }
        for (int pv = 0; pv < c.nv; pv++) {
            if (c.w[pv] != 0.0f) continue;
            prims.push_back((int)idx.size());
            prims.push_back(1);
            idx.push_back(pv);
            for (int k = 0; k < 3; k++) rest_pos.push_back(c.pos0[3 * pv + k]);
            type.push_back(PBD_ANCHOR);
            stiff.push_back(-1.0f);
        }
        Writer w(constr_path, digits);
        w.ints("Indices", idx.data(), idx.size());
        w.float3s("P", c.pos0.data(), c.nv);
        w.int2s("primList", prims.data(), prims.size() / 2);
        w.float3s("prdPBuffer", p, c.nv);
        w.float3s("restPosBuffer", rest_pos.data(), rest_pos.size() / 3);
        w.floats("restLengthBuffer", rest_len.data(), rest_len.size());
        w.floats("stiffnessBuffer", stiff.data(), stiff.size());
        w.ints("constraintType", type.data(), type.size());
        if (c.nb > 0) w.floats("restAngleBuffer", rest_ang.data(), rest_ang.size());
        w.finish(constr_path);
    }
}

// Fills a fresh Cloth whose prm holds the library defaults.
void read_cache(const char* mesh_path, const char* constr_path, Cloth& c) {
    Lines M = read_lines(mesh_path);
    PBD_CHECK(has(M, "P"), PBD_ERR_IO, std::string("no positions (P) in ") + mesh_path);
    std::vector<float> x = get_vecf(M, "P", 3);
    int nv = (int)(x.size() / 3);
    std::vector<int> ind = get_ints(M, "Indices");
    std::vector<int> mprims = get_veci(M, "primList", 2);
    std::vector<int> tris;
    if (mprims.empty()) {
        PBD_CHECK(ind.size() % 3 == 0, PBD_ERR_IO, "mesh index count is not a multiple of 3");
        tris = ind;
    } else {
        for (size_t t = 0; t < mprims.size() / 2; t++) {
            int s = mprims[2 * t], n = mprims[2 * t + 1];
            PBD_CHECK(n == 3, PBD_ERR_IO, "mesh prims must be triangles");
            PBD_CHECK(s >= 0 && (size_t)s + 3 <= ind.size(), PBD_ERR_IO, "mesh prim outside the index list");
            tris.insert(tris.end(), ind.begin() + s, ind.begin() + s + 3);
        }
    }
    for (int i : tris) PBD_CHECK(i >= 0 && i < nv, PBD_ERR_IO, "triangle index out of range");
    std::vector<float> vel = get_vecf(M, "velBuffer", 3);
    PBD_CHECK(vel.empty() || vel.size() == x.size(), PBD_ERR_IO, "velBuffer count differs from P");
    std::vector<float> mass = get_floats(M, "massBuffer");
    PBD_CHECK(mass.empty() || (int)mass.size() == nv, PBD_ERR_IO, "massBuffer count differs from P");
    for (float m : mass) PBD_CHECK(m >= 0.0f, PBD_ERR_IO, "negative massBuffer weight");

    c.nv = nv;
    c.nt = (int)tris.size() / 3;
    c.tris = std::move(tris);
    c.pos0 = x;
    c.x_start = std::move(x);
    c.v_start = std::move(vel);
    c.w_in = std::move(mass);  // a weight of 0 pins its vertex
    pbd_cloth_params& prm = c.prm;
    if (has(M, "dampingRate")) {
        std::vector<float> d = get_floats(M, "dampingRate");
        PBD_CHECK(d.size() == 1 && d[0] > 0.0f, PBD_ERR_IO, "dampingRate must be one positive number");
        prm.damping = d[0];
    }
    if (has(M, "gravity")) {
        std::vector<float> g = get_vecf(M, "gravity", 3);
        PBD_CHECK(g.size() == 3, PBD_ERR_IO, "gravity must be one float3");
        for (int k = 0; k < 3; k++) prm.gravity[k] = g[k];
    }
    if (!constr_path || !constr_path[0]) return;

    Lines K = read_lines(constr_path);
    std::vector<float> cp = get_vecf(K, "P", 3);
    PBD_CHECK(cp.empty() || cp.size() == c.pos0.size(), PBD_ERR_IO, "constraint P count differs from the mesh");
    if (!cp.empty()) c.pos0 = std::move(cp);
    std::vector<float> prd = get_vecf(K, "prdPBuffer", 3);
    PBD_CHECK(prd.empty() || prd.size() == c.pos0.size(), PBD_ERR_IO, "prdPBuffer count differs from the mesh");
    c.p_start = std::move(prd);
    std::vector<int> kind = get_ints(K, "Indices");
    std::vector<int> kprims = get_veci(K, "primList", 2);
    std::vector<int> type = get_ints(K, "constraintType");
    std::vector<float> stiff = get_floats(K, "stiffnessBuffer");
    std::vector<float> rl = get_floats(K, "restLengthBuffer");
    std::vector<float> ra = get_floats(K, "restAngleBuffer");
    size_t np = kprims.size() / 2;
    PBD_CHECK(type.size() == np, PBD_ERR_IO, "constraintType count differs from primList");
    PBD_CHECK(stiff.size() == np, PBD_ERR_IO, "stiffnessBuffer count differs from primList");
    size_t kd = 0, kb = 0;
    float sd = 0.0f, sb = 0.0f;
    std::vector<float> rest_b;
    auto same_stiffness = [&](float& seen, size_t count, float s, const char* kind_name) {
        PBD_CHECK(s >= 0.0f && s <= 1.0f, PBD_ERR_IO, std::string(kind_name) + " stiffness must be in [0,1]");
        PBD_CHECK(count == 0 || seen == s, PBD_ERR_IO,
                  std::string("stiffness differs between ") + kind_name +
                      " prims (one value per constraint type is supported)");
        seen = s;
    };
    for (size_t i = 0; i < np; i++) {
        int s = kprims[2 * i], n = kprims[2 * i + 1];
        PBD_CHECK(s >= 0 && n >= 1 && (size_t)s + n <= kind.size(), PBD_ERR_IO, "prim outside the index list");
        const int* q = &kind[s];
        for (int k = 0; k < n; k++) PBD_CHECK(q[k] >= 0 && q[k] < nv, PBD_ERR_IO, "prim vertex out of range");
        if (type[i] == PBD_DISTANCE) {
            PBD_CHECK(n == 2, PBD_ERR_IO, "a DISTANCE prim must have 2 vertices");
            PBD_CHECK(kd < rl.size(), PBD_ERR_IO, "fewer rest lengths than DISTANCE prims");
            same_stiffness(sd, kd, stiff[i], "DISTANCE");
            c.user_pairs.insert(c.user_pairs.end(), q, q + 2);
            c.user_rest.push_back(rl[kd++]);
        } else if (type[i] == PBD_BENDING) {
            PBD_CHECK(n == 4, PBD_ERR_IO, "a BENDING prim must have 4 vertices");
            PBD_CHECK(kb < ra.size(), PBD_ERR_IO, "fewer rest angles than BENDING prims");
            same_stiffness(sb, kb, stiff[i], "BENDING");
            c.user_quads.insert(c.user_quads.end(), q, q + 4);
            rest_b.push_back(ra[kb++]);
        } else if (type[i] == PBD_ANCHOR) {
            PBD_CHECK(n == 1, PBD_ERR_IO, "an ANCHOR prim must have 1 vertex");
            if (std::find(c.pins.begin(), c.pins.end(), q[0]) == c.pins.end()) c.pins.push_back(q[0]);
        } else {
            throw Error(PBD_ERR_IO, "unknown constraintType value");
        }
    }
    PBD_CHECK(kd == rl.size(), PBD_ERR_IO, "more rest lengths than DISTANCE prims");
    PBD_CHECK(kb == ra.size(), PBD_ERR_IO, "more rest angles than BENDING prims");
    c.has_user_pairs = true;
    c.has_user_quads = true;
    c.ct = (kd ? PBD_DISTANCE : 0u) | (kb ? PBD_BENDING : 0u) | (c.pins.empty() ? 0u : PBD_ANCHOR);
    if (kd) c.stiff[0] = sd;
    if (kb) c.stiff[1] = sb;
    c.user_rest.insert(c.user_rest.end(), rest_b.begin(), rest_b.end());
}

}  // namespace pbd
