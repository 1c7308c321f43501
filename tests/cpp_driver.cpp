// A C++ caller of libpbd.so: the reference's driver loop
// (PositionBasedDynamics.cpp:256-295 -- PBDObject + SolverPBD, then per frame
// and substep Advect / ProjectConstraint(GAUSSSEIDEL, iterations) /
// Integration) written against include/pbd.h.
//
//   cpp_driver OUT.bin [substeps]
//
// A 32 x 32 cloth (1 m, pins at the two top corners, jitter
// 1e-4 m from seed 0, dt 1/600 s, 20 sweeps) is stepped twice from the same
// start: with pbd_step (substeps replayed from a captured graph) and with the
// split loop pbd_advect / pbd_project / pbd_integrate, the reference's three
// calls.  The two must agree bit for bit; the program writes x then v of the
// first (float32, 3 per vertex) to OUT.bin and exits 0, or prints the error
// and exits 1.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "pbd.h"

namespace {

bool ok(int code, const char* what) {
    if (code == PBD_OK) return true;
    std::fprintf(stderr, "%s failed (%d): %s\n", what, code, pbd_last_error());
    return false;
}

#define CHECK(call)                   \
    do {                              \
        if (!ok((call), #call)) return false; \
    } while (0)

constexpr int kRes = 32;
constexpr float kDt = 1.0f / 600.0f;
constexpr int kIterations = 20;

// Builds the 32 x 32 cloth in a new context (PBDObject(...).Init() and
// InitAnchorConstr, PBD_cuda.cu:125-153,831-869).
bool make_cloth(pbd_ctx** ctx, int* cloth, int graph) {
    CHECK(pbd_create(0, ctx));
    pbd_grid_desc g{};
    g.res_x = kRes;
    g.res_y = kRes;
    g.size_x = 1.0f;
    g.size_y = 1.0f;
    g.jitter = 1e-4f;
    g.seed = 0;
    CHECK(pbd_cloth_create_grid(*ctx, &g, cloth));
    const int pins[2] = {0, kRes - 1};
    CHECK(pbd_cloth_set_pins(*ctx, *cloth, pins, 2));
    CHECK(pbd_init(*ctx, 20000));
    pbd_solver_params sp{};
    sp.dt = kDt;
    sp.iterations = kIterations;
    sp.tol = -1.0;
    sp.check_every = 1;
    sp.path = PBD_PATH_AUTO;
    sp.graph = graph;
    CHECK(pbd_set_solver_params(*ctx, &sp));
    return true;
}

bool run(const char* out_path, int substeps) {
    pbd_ctx* a = nullptr;
    pbd_ctx* b = nullptr;
    int ca = 0, cb = 0;
    const int nv = kRes * kRes;
    std::vector<float> xa(3 * nv), va(3 * nv), xb(3 * nv), vb(3 * nv);
    bool good = make_cloth(&a, &ca, 1) && make_cloth(&b, &cb, 0);
    if (good) {
        // one pbd_step call per frame of 10 substeps, as the reference's
        // frame loop; the remainder in a last call
        for (int done = 0; good && done < substeps; done += 10)
            good = ok(pbd_step(a, substeps - done < 10 ? substeps - done : 10), "pbd_step");
        for (int s = 0; good && s < substeps; s++)
            good = ok(pbd_advect(b, kDt), "pbd_advect") && ok(pbd_project(b, kIterations), "pbd_project") &&
                   ok(pbd_integrate(b, kDt), "pbd_integrate");
    }
    good = good && ok(pbd_cloth_get_state(a, ca, PBD_STATE_X, xa.data()), "get x") &&
           ok(pbd_cloth_get_state(a, ca, PBD_STATE_V, va.data()), "get v") &&
           ok(pbd_cloth_get_state(b, cb, PBD_STATE_X, xb.data()), "get x") &&
           ok(pbd_cloth_get_state(b, cb, PBD_STATE_V, vb.data()), "get v");
    if (a) good = ok(pbd_destroy(a), "pbd_destroy") && good;
    if (b) good = ok(pbd_destroy(b), "pbd_destroy") && good;
    if (!good) return false;
    if (std::memcmp(xa.data(), xb.data(), sizeof(float) * xa.size()) != 0 ||
        std::memcmp(va.data(), vb.data(), sizeof(float) * va.size()) != 0) {
        std::fprintf(stderr, "pbd_step and the advect / project / integrate loop differ\n");
        return false;
    }
    FILE* f = std::fopen(out_path, "wb");
    if (!f) {
        std::fprintf(stderr, "cannot write %s\n", out_path);
        return false;
    }
    const bool wrote = std::fwrite(xa.data(), sizeof(float), xa.size(), f) == xa.size() &&
                       std::fwrite(va.data(), sizeof(float), va.size(), f) == va.size();
    return std::fclose(f) == 0 && wrote;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s OUT.bin [substeps]\n", argv[0]);
        return 2;
    }
    const int substeps = argc > 2 ? std::atoi(argv[2]) : 10;
    if (substeps < 1) {
        std::fprintf(stderr, "substeps must be positive\n");
        return 2;
    }
    if (!run(argv[1], substeps)) return 1;
    std::printf("cpp_driver: %d substeps of the %dx%d cloth written to %s\n", substeps, kRes, kRes, argv[1]);
    return 0;
}
