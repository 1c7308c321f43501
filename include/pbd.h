/*
 * pbd.h -- C ABI of the MI355X-native position-based-dynamics cloth solver.
 *
 * One shared library (libpbd.so) exports these functions; plain pointers and
 * sizes only, no C++ or torch types.  Host arrays are caller-owned and copied;
 * device memory is owned by the context.  Every function returns PBD_OK (0)
 * or a negative PBD_ERR_* code, and pbd_last_error() describes the last
 * failure of the calling thread.  A context uses one HIP stream and must be
 * driven by one host thread at a time.
 *
 * The reference (vivienldy/PositionBasedDynamicsOld) exposes this path as
 * C++ classes only; each entry point below names the member it replaces.
 * Paths are relative to PositionBasedDynamicsCUDA/PositionBasedDynamicsCUDA/.
 */
#ifndef PBD_H
#define PBD_H

#include <stdint.h>

#if defined(__GNUC__)
#define PBD_API __attribute__((visibility("default")))
#else
#define PBD_API
#endif

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pbd_ctx pbd_ctx;

enum {
    PBD_OK = 0,
    PBD_ERR_ARG = -1,       /* bad argument (null pointer, index or size out of range) */
    PBD_ERR_STATE = -2,     /* call out of order (e.g. step before init) */
    PBD_ERR_HIP = -3,       /* HIP runtime error */
    PBD_ERR_LIMIT = -4,     /* a size limit of the implementation was exceeded */
    PBD_ERR_IO = -5,        /* file could not be read or parsed */
    PBD_ERR_COLOR = -6      /* colouring did not converge in the allowed rounds */
};

/* ConstraintType values, PBD_Basic.cuh:38-43. */
enum { PBD_DISTANCE = 0x001, PBD_BENDING = 0x010, PBD_ANCHOR = 0x100 };

/* Which state buffer pbd_cloth_get_state / pbd_cloth_set_state address. */
enum { PBD_STATE_X = 0, PBD_STATE_V = 1, PBD_STATE_P = 2 };

/* Projection path. */
enum {
    PBD_PATH_AUTO = 0,      /* LDS tiles (whole-cloth or dependency-cone tiles) */
    PBD_PATH_GLOBAL = 1,    /* one launch per colour over global memory */
    PBD_PATH_JACOBI = 2     /* colour-free Jacobi with constraint averaging, one launch per sweep
                               (SolverType::JACOBI, declared but not implemented, PBD_Basic.cuh:46-51) */
};

/* Procedural grid cloth (PBDObject(damping, g, resX, resY, sizeX, sizeY, ht)
 * + PBDObject::Init(), PBD_Basic.cuh:155-158, PBD_cuda.cu:644-648,831-869). */
typedef struct {
    int res_x, res_y;
    float size_x, size_y;
    float origin[3];        /* translation added after generation (reference: 0) */
    float jitter;           /* uniform +-jitter per coordinate from `seed` (0 = none) */
    uint64_t seed;
} pbd_grid_desc;

/* Per-cloth physical parameters (PBDObject::dampingRate/gravity,
 * SolverPBD sphere, PBD_Basic.cuh:190-192,242-244). */
typedef struct {
    float damping;          /* dampingRate; velocity scaled by damping^dt each substep */
    float gravity[3];
    float wind[3];          /* wind velocity (m/s) */
    float wind_drag;        /* drag coefficient (1/s) toward the wind velocity; 0 = off */
    int sphere_on;
    float sphere_c[3];
    float sphere_r;
    int box_on;
    float box_lo[3];
    float box_hi[3];
} pbd_cloth_params;

/* Solver parameters (SolverPBD::ProjectConstraint(st, iterations), PBD_Basic.cuh:226). */
typedef struct {
    float dt;               /* substep length */
    int iterations;         /* sweeps per substep (tol < 0) or sweep cap (tol >= 0) */
    double tol;             /* stop when max |d-L|/L <= tol after a check; < 0 = fixed sweeps */
    int check_every;        /* sweeps between residual checks (chunk length) */
    int path;               /* PBD_PATH_* */
    int tile_sweeps;        /* sweeps per launch for multi-tile cloths (0 = check_every) */
    int tile_max_verts;     /* LDS vertices per tile (0 = from the device's LDS size) */
    int block_threads;      /* threads per workgroup of the tile kernel (0 = 1024) */
    int graph;              /* 1: pbd_step replays each substep from a captured hipGraph whenever the
                               substep makes no host decision (fixed sweeps, no collisions, profiling off) */
    int warm_iterations;    /* with the warm start on: sweep budget of a cloth whose prediction was kept
                               (0 = iterations); a cloth whose prediction is discarded runs `iterations`,
                               exactly the plain solver's substep */
    int warm_rule;          /* fallback rule: the prediction is kept iff its residual does not exceed the
                               unwarmed start's, measured as max |d-L|/L (0) or as the RMS of |d-L|/L (1) */
    float warm_gate;        /* a vertex whose one-ring distance prims all have |d-L|/L < warm_gate gets
                               no predicted correction (0 = every vertex gets one); a cloth with no strained
                               vertex then keeps the zero prediction and runs warm_iterations sweeps */
    float jacobi_omega;     /* PBD_PATH_JACOBI: a vertex moves by omega / n times the sum of the corrections
                               its n prims give it (0 = 1; 1 to 2 over-relaxes) */
    int split_whole;        /* 1: a cloth that fits one LDS tile is stepped tile_sweeps sweeps per launch like a
                               multi-tile cloth (its one tile loads the whole cloth, no halo) instead of all its
                               sweeps in one launch, so that no launch waits on one cloth's whole budget
                               (0 = one launch; results are identical either way) */
    int per_chunk_launches; /* under per-cloth budgets (the warm start) with fixed sweeps: 0 = the substep's
                               tile work is one launch whose workgroups take (chunk, tile) items from a queue
                               built on the device after a residual kernel has made the keep decision, a
                               tile's later chunk waiting only for its neighbour tiles' previous chunk;
                               1 = one launch per chunk of tile_sweeps sweeps; 2 = as 0, but in a warm
                               substep the queued launch computes the fallback residuals itself and keeps or
                               discards each cloth's prediction (no residual kernel; a cloth's tiles wait for
                               each other, and later chunks are appended once a cloth is decided; when a
                               multi-tile cloth has as many tiles as the GPU holds workgroups, as 0).
                               Results are identical in all three; 0 measured fastest */
} pbd_solver_params;

typedef struct {
    int nv, nt;             /* vertices, triangles */
    int ne, nb;             /* distance prims, bending prims */
    int nprims;             /* ne + nb (coloured prim set) */
    int ncolors;
    int color_rounds;       /* assign/resolve rounds the colouring took */
    int adj_len;            /* total prim->prims neighbour entries */
    int ntiles;             /* LDS tiles planned for this cloth */
    int whole;              /* 1 if the cloth is one LDS-resident tile */
    int max_loaded;         /* largest loaded vertex set over its tiles */
} pbd_cloth_info;

/* ---- lifecycle ---------------------------------------------------------- */
PBD_API const char* pbd_last_error(void);
PBD_API int pbd_create(int device, pbd_ctx** out);
PBD_API int pbd_destroy(pbd_ctx* ctx);
PBD_API int pbd_synchronize(pbd_ctx* ctx);

/* ---- model build (PBDObject / ConstraintPBD) ------------------------------ */
/* PBDObject(...,resX,resY,sizeX,sizeY) + initMeshTopol (PBD_cuda.cu:675-705,831-869) */
PBD_API int pbd_cloth_create_grid(pbd_ctx* ctx, const pbd_grid_desc* desc, int* cloth);
/* PBDObject(damping, g, ht) + IO::ReadTopolFromTxt (PBD_Basic.cuh:532-601), from memory */
PBD_API int pbd_cloth_create_mesh(pbd_ctx* ctx, const float* pos, int nv, const int* tris, int nt, int* cloth);
/* PBDObject::Init(topolFile, distConstrFile) (PBD_cuda.cu:650-658): the reference
 * text formats, IO::ReadTopolFromTxt + IO::ReadDistConstrFromTxt (PBD_Basic.cuh:532-688) */
PBD_API int pbd_cloth_load_txt(pbd_ctx* ctx, const char* topol_path, const char* constr_path, int* cloth);
/* PBDObject::ContinueSimInit(meshTopolPath, constrPath) (PBD_cuda.cu:660-667):
 * IO::ReadTopolFromCache + IO::ReadConstraintFromCache (PBD_Basic.cuh:326-530).
 * The cloth starts from the cached P, velBuffer and prdPBuffer, with massBuffer
 * as its weights (0 pins a vertex), dampingRate and gravity, and the cached
 * prims (DISTANCE pairs, BENDING quads, ANCHOR pins) with their cached rest
 * values.  One stiffness per constraint type; constr_path may be NULL (mesh
 * edges from the triangles, rest values from P). */
PBD_API int pbd_cloth_load_cache(pbd_ctx* ctx, const char* mesh_path, const char* constr_path, int* cloth);
/* PBDObject::SaveMeshTopol + PBDObject::SaveConstraint (PBD_cuda.cu:907-938):
 * the cloth's current state and constraints in the reference's cache format
 * (after pbd_init).  precision = significant digits of every float (0 = 6, the
 * reference's stream default; 9 makes the round trip exact).  Either path may
 * be NULL to skip that file. */
PBD_API int pbd_cloth_save_cache(pbd_ctx* ctx, int cloth, const char* mesh_path, const char* constr_path,
                                 int precision);
/* Explicit distance prims (IO::ReadDistConstrFromTxt pairs, PBD_Basic.cuh:654-675);
 * without it a grid cloth uses InitDistanceIndices order (PBD_cuda.cu:19-54) and a
 * mesh cloth uses its triangle edges in first-sighting order. */
PBD_API int pbd_cloth_set_distance_pairs(pbd_ctx* ctx, int cloth, const int* pairs, int ne);
/* PBDObject::SetConstrOption(ct, stiffness[DISTANCE, BENDING, ANCHOR]) (PBD_cuda.cu:669-673) */
PBD_API int pbd_cloth_set_constraint_option(pbd_ctx* ctx, int cloth, unsigned ct, const float* stiffness3);
/* Pins (replaces InitAnchorConstr, PBD_cuda.cu:125-153): weight 0 for the listed vertices */
PBD_API int pbd_cloth_set_pins(pbd_ctx* ctx, int cloth, const int* verts, int n);
PBD_API int pbd_cloth_set_params(pbd_ctx* ctx, int cloth, const pbd_cloth_params* params);
/* The cloth's current parameters (PBDObject::dampingRate/gravity, PBD_Basic.cuh:190-192),
 * e.g. those a cache file set, to change one field with pbd_cloth_set_params. */
PBD_API int pbd_cloth_get_params(pbd_ctx* ctx, int cloth, pbd_cloth_params* params);
/* initConstr: constraints, GenePoint2PrimsMap/GenePrim2PrimsMap, InitGPUBuffers,
 * EdgeColoring(color_rounds) on the GPU, SortEdgesColors + EvalWorksets
 * (PBD_cuda.cu:798-829,155-202,707-750,487-625), for every cloth created so far. */
PBD_API int pbd_init(pbd_ctx* ctx, int color_rounds);
PBD_API int pbd_cloth_get_info(pbd_ctx* ctx, int cloth, pbd_cloth_info* info);
/* ConstraintPBD::topol.primList/indices + restLengthBuffer/rest angle (PBD_Basic.cuh:97-108) */
PBD_API int pbd_cloth_get_prims(pbd_ctx* ctx, int cloth, int* prim_list, int* indices, float* rest);
/* ConstraintPBD::Prim2PrimsMap (startNumList, indices) (PBD_cuda.cu:170-202) */
PBD_API int pbd_cloth_get_adjacency(pbd_ctx* ctx, int cloth, int* start_num, int* adj);
/* ConstraintPBD::color, sortedPrimId, colorWorksets (PBD_cuda.cu:546-625) */
PBD_API int pbd_cloth_get_colors(pbd_ctx* ctx, int cloth, int* color, int* sorted_prim, int* worksets);

/* ---- solver (SolverPBD) ---------------------------------------------------- */
PBD_API int pbd_set_solver_params(pbd_ctx* ctx, const pbd_solver_params* params);
/* SolverPBD::Advect(dt) (PBD_cuda.cu:981-1052) */
PBD_API int pbd_advect(pbd_ctx* ctx, float dt);
/* SolverPBD::ProjectConstraint(GAUSSSEIDEL, iterations) (PBD_cuda.cu:1054-1070,1308-1435) */
PBD_API int pbd_project(pbd_ctx* ctx, int iterations);
/* SolverPBD::Integration(dt) (PBD_cuda.cu:1437-1518) */
PBD_API int pbd_integrate(pbd_ctx* ctx, float dt);
/* The reference's main-loop body Advect -> ProjectConstraint -> Integration
 * (PositionBasedDynamics.cpp:263-290), `substeps` times, fused. */
PBD_API int pbd_step(pbd_ctx* ctx, int substeps);
/* SolverPBD::SetTarget(obj) and the same loop body on that one object
 * (PBD_Basic.cuh:224-233, PositionBasedDynamics.cpp:263-290): substeps of a
 * single cloth whose launches cover only its own vertices and prims -- the
 * unbatched solver.  Collisions and the warm start run on the whole batch
 * only (PBD_ERR_STATE while enabled). */
PBD_API int pbd_step_cloth(pbd_ctx* ctx, int cloth, int substeps);
PBD_API int pbd_cloth_get_state(pbd_ctx* ctx, int cloth, int which, float* xyz);
PBD_API int pbd_cloth_set_state(pbd_ctx* ctx, int cloth, int which, const float* xyz);
/* Convergence residual over distance prims: max and RMS of |d - L| / L (FP64). */
PBD_API int pbd_cloth_residual(pbd_ctx* ctx, int cloth, double* max_rel, double* rms);
/* Residual of every cloth at once (arrays of the cloth count). */
PBD_API int pbd_residual_all(pbd_ctx* ctx, double* max_rel, double* rms);
/* Exact binary checkpoint of the whole context (x and v of every vertex in
 * packed order, 3 floats each); replaces the lossy text caches of
 * PBDObject::SaveMeshTopol / ContinueSimInit (PBD_cuda.cu:660-667,907-925). */
PBD_API int pbd_num_vertices(pbd_ctx* ctx, int* nv);
PBD_API int pbd_state_save(pbd_ctx* ctx, float* x, float* v);
PBD_API int pbd_state_load(pbd_ctx* ctx, const float* x, const float* v);
/* Device-side copy of x and v (restore = 0) and its restore (restore = 1). */
PBD_API int pbd_snapshot(pbd_ctx* ctx, int restore);
/* Residual tolerance of one cloth, overriding pbd_solver_params.tol (< 0 restores it). */
PBD_API int pbd_cloth_set_tolerance(pbd_ctx* ctx, int cloth, double tol);
/* Sweeps the last projection of this cloth ran. */
PBD_API int pbd_cloth_last_sweeps(pbd_ctx* ctx, int cloth, int* sweeps);

/* ---- broadphase + CCD ---------------------------------------------------- */
/* SpatialHashSystem(pos, tris, ht, gridCenter, gridSize, cellSize) grid
 * (SpatialHashSystem.h:15, SpatialHashSystem.cpp:49-66,88-98).  mode 0 is the
 * reference's dense grid (key (z*gy+y)*gx+x, -1 outside); mode 1 hashes cell
 * coordinates into a 2^table_bits table (every cell has a slot). */
typedef struct {
    int mode;
    float centre[3];
    int size[3];
    float cell[3];
    int table_bits;
} pbd_hash_desc;

/* SpatialHashSystem::InitSH + UpdateSH on caller positions
 * (SpatialHashSystem.cpp:88-121): keys, stable sort, cell ranges. */
PBD_API int pbd_sh_build(pbd_ctx* ctx, const pbd_hash_desc* desc, const float* pos, int nv, const int* tris,
                         int nt);
/* pbd_sh_build for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * It has no counterpart in the reference.  Every pointer is device memory of
 * `device`: pos (3 floats per vertex), tris (3 ints per triangle) in; out the
 * sorted keys (-1 first) and triangle ids and the unsorted keys (nt ints
 * each), and the cell ranges (pbd_hash_desc cells, or 2^table_bits slots,
 * ints each), as pbd_sh_get returns them.  The work is queued on `stream`
 * (a hipStream_t; NULL is the null stream) and the call returns without
 * waiting; the outputs are ready when the stream reaches that point.  It
 * allocates nothing: the sort's scratch is one static device array of
 * 8.5 MiB per GPU, held until the process ends, which limits nt to 2^20
 * (PBD_ERR_LIMIT above).  Calls that overlap on one GPU race on that array.
 * A triangle with a vertex index outside [0, nv) is keyed -1 (outside). */
PBD_API int pbd_sh_build_noalloc(int device, void* stream, const pbd_hash_desc* desc, const float* pos, int nv,
                                 const int* tris, int nt, int* sorted_keys, int* sorted_ids, int* raw_keys,
                                 int* cell_start, int* cell_end);
PBD_API int pbd_sh_num_cells(pbd_ctx* ctx, int* ncells);
/* m_hTriangleHash/m_hTriangleId after sortTrianglesCPU, the unsorted keys of
 * calcHashCPU, and m_hCellStart/m_hCellEnd (SpatialHashSystem.cpp:300-389). */
PBD_API int pbd_sh_get(pbd_ctx* ctx, int* sorted_keys, int* sorted_ids, int* raw_keys, int* cell_start,
                       int* cell_end);
/* FindNeighbors over every triangle as CCD_SH_Extended consumes it
 * (SpatialHashSystem.cpp:134-187, CCD.cpp:1541-1586): the set of candidate
 * (vertex, face) pairs, v not in f, as sorted keys (v << 32 | f). */
PBD_API int pbd_sh_pairs(pbd_ctx* ctx, long long cap, int64_t* keys, long long* count);
/* CCD_SH_Extended(contactData, shs, meshTopol, prdPBuffer, thickness)
 * (CCD.cpp:1522-1587): x = start-of-step positions, p = predicted. */
PBD_API int pbd_ccd_detect(pbd_ctx* ctx, const pbd_hash_desc* desc, const float* x, const float* p, int nv,
                           const int* tris, int nt, float thickness, long long* ncontacts);
/* Detection with the swept-box candidate set (broadphase 1 of
 * pbd_collision_params); contacts equal CCD_N2's (CCD.cpp:1040-1098) over
 * all (v, f), v a vertex of some triangle and not in f.  No reference
 * counterpart for the candidate set itself. */
PBD_API int pbd_ccd_detect_swept(pbd_ctx* ctx, const pbd_hash_desc* desc, const float* x, const float* p, int nv,
                                 const int* tris, int nt, float thickness, long long* npairs, long long* ncontacts);
/* Candidate pairs of the last detection, ascending keys (v << 32 | f). */
PBD_API int pbd_ccd_get_pairs(pbd_ctx* ctx, long long cap, int64_t* keys, long long* count);
/* ContactData (CCD_Basic.h:12-31) of the last detection, ascending (v, f). */
PBD_API int pbd_ccd_get_contacts(pbd_ctx* ctx, long long cap, int* vf, float* t, float* n, double* w);
/* Edge-edge contacts (Contact::EE, declared but not implemented in the
 * reference: CCD_Basic.h:14, CCD.cpp:925-928), added to the vertex-face
 * contacts of a pbd_ccd_detect / pbd_ccd_detect_swept call on the same
 * scene and settings: the same x, p and triangles, byte for byte, the same
 * thickness and the same grid description (none otherwise).  Candidates: every edge pair of the triangle
 * pairs pbd_ccd_detect visits (centroid cells of desc) that shares no vertex
 * and has no vertex that is a mesh neighbour of the other edge's.
 * Test: the first time in [0, 1] (coplanarity roots of the four points, then
 * 0 and 1) at which the segments come within 2 x thickness, for edges that
 * start at least one thickness apart along their normal. */
PBD_API int pbd_ccd_detect_ee(pbd_ctx* ctx, const pbd_hash_desc* desc, const float* x, const float* p, int nv,
                              const int* tris, int nt, float thickness, long long* npairs, long long* ncontacts);
/* Edge-edge candidates of the last pbd_ccd_detect_ee: 4 vertex ids each
 * (a0 < a1, b0 < b1), ordered by edge id pair. */
PBD_API int pbd_ccd_get_ee_pairs(pbd_ctx* ctx, long long cap, int* quads, long long* count);
/* Edge-edge contacts of the last pbd_ccd_detect_ee: vertices (a0, a1, b0, b1),
 * time t, end-state edge normal n (3), closest-point weights
 * w = (1 - s, s, 1 - u, u). */
PBD_API int pbd_ccd_get_ee_contacts(pbd_ctx* ctx, long long cap, int* quads, float* t, float* n, double* w,
                                    long long* count);
/* CollisionResolveNew(meshTopol, prdPBuffer, contactData, iterations, thickness)
 * (CCD.cpp:1874-1994) on the last detection, Jacobi form; p is updated in
 * place, w[v] == 0 marks static vertices.  Edge-edge contacts of a
 * following pbd_ccd_detect_ee are resolved in the same passes. */
PBD_API int pbd_ccd_resolve(pbd_ctx* ctx, float* p, const float* w, float thickness, int passes);

/* In-solver collision (SolverPBD::ProjectConstraintWithColli, PBD_cuda.cu:1130-1210):
 * detection every `detect_every` sweeps and at the last sweep, contact
 * response after every sweep, over all cloths of the context. */
typedef struct {
    int enabled;
    pbd_hash_desc grid;      /* mode 1: centre ignored, cell from cell_scale when > 0 */
    float cell_scale;        /* cell = cell_scale x mean rest length (0 = use grid.cell) */
    float thickness;
    int resolve_passes;
    int detect_every;
    /* The response skips a face whose unnormalised normal has
     * |(p2 - p1) x (p3 - p1)|^2 <= min_normal2 (m^4).  0 selects the
     * reference's fixed 1e-6 (VFResolveTest, CCD.cpp:2008), which also skips
     * every well-shaped face with edges below about 3.4 cm. */
    float min_normal2;
    /* Candidate pairs: 0 = the reference's centroid cells (FindNeighbors,
     * SpatialHashSystem.cpp:134-187), which misses faces that extend more
     * than about a cell from their centroid; 1 = swept boxes, whose contacts
     * equal the all-pairs test (CCD_N2, CCD.cpp:1040-1098) at any cell size
     * (the cell only sets the cost).  No reference counterpart for 1. */
    int broadphase;
    /* 1: edge-edge contacts too (pbd_ccd_detect_ee's test and response),
     * from the centroid-cell triangle pairs; no reference counterpart
     * (Contact::EE is declared but not implemented, CCD_Basic.h:14). */
    int edge_edge;
} pbd_collision_params;
PBD_API int pbd_set_collision(pbd_ctx* ctx, const pbd_collision_params* params);
PBD_API int pbd_last_contacts(pbd_ctx* ctx, long long* ncontacts, long long* npairs);

/* ---- learned warm start (no counterpart in the reference) -------------------
 * A per-vertex MLP (22 -> 64 -> 64 -> 3, GELU, BF16 weights on MFMA, FP32
 * accumulation) predicts the Gauss-Seidel correction of each vertex from its
 * pre-projection state; the prediction is kept for a cloth only when it does
 * not raise that cloth's residual above the unwarmed start. */
PBD_API int pbd_ws_num_params(int* n);
PBD_API int pbd_ws_init(pbd_ctx* ctx, uint64_t seed);
PBD_API int pbd_ws_get_params(pbd_ctx* ctx, float* params);
PBD_API int pbd_ws_set_params(pbd_ctx* ctx, const float* params);
/* Records (features, correction of `sweeps` sweeps) for every vertex over
 * `substeps` plain substeps of the current scene.  With collisions enabled
 * (pbd_set_collision) each recorded substep runs the collision sweeps, so the
 * correction includes the contact response. */
PBD_API int pbd_ws_collect(pbd_ctx* ctx, int sweeps, int substeps);
PBD_API int pbd_ws_num_samples(pbd_ctx* ctx, long long* n);
PBD_API int pbd_ws_get_samples(pbd_ctx* ctx, long long first, long long n, float* X, float* Y);
PBD_API int pbd_ws_set_samples(pbd_ctx* ctx, const float* X, const float* Y, long long n);
PBD_API int pbd_ws_forward(pbd_ctx* ctx, const float* X, int n, float* Y);
PBD_API int pbd_ws_gradient(pbd_ctx* ctx, long long first, int n, float* grad, double* loss);
PBD_API int pbd_ws_adam_step(pbd_ctx* ctx, const float* grad, float lr);
PBD_API int pbd_ws_train(pbd_ctx* ctx, int epochs, int batch, float lr, uint64_t seed, double* last_loss);
/* Feature mean[22], feature std[22], target std[3]; set by the first
 * pbd_ws_train from the collected samples. */
PBD_API int pbd_ws_get_normalization(pbd_ctx* ctx, float* norm47);
PBD_API int pbd_ws_set_normalization(pbd_ctx* ctx, const float* norm47);
PBD_API int pbd_ws_enable(pbd_ctx* ctx, int on);
PBD_API int pbd_ws_last_used(pbd_ctx* ctx, int cloth, int* used);
/* The correction the network predicted for this cloth in the last warm-started
 * substep (3 floats per vertex, zero for pinned vertices), before the
 * fallback decision. */
PBD_API int pbd_ws_last_delta(pbd_ctx* ctx, int cloth, float* xyz);

/* ---- measurement (no counterpart in the reference, whose only timing is the
 * host Timer around calls, timer.cpp:10-20) --------------------------------- */
typedef struct {
    double tile_ms;          /* summed duration of LDS-tile kernel launches (HIP events on the context stream) */
    long long tile_launches;
    double tile_alg_bytes;   /* per-sweep model bytes of those launches (68 B/distance prim/sweep,
                                132 B/bending prim/sweep, 100 B/vertex for fused advect+integrate);
                                the launches read each loaded vertex once for several sweeps, so
                                tile_bytes below is what they move */
    double step_ms;          /* summed device-synchronised wall time of pbd_step calls */
    long long substeps;
    long long vertex_sweeps; /* sum over cloths and substeps of V x sweeps */
    long long sweeps_max;    /* largest per-cloth sweep count seen */
    long long graph_replays; /* substeps replayed from a captured graph */
    /* contact detection of the in-solver collision path, summed over detections */
    long long ccd_detections;
    long long ccd_triangles;    /* triangles hashed */
    long long ccd_cell_entries; /* (cell, face) entries (swept-box candidates; one per triangle otherwise) */
    long long ccd_raw_pairs;    /* candidate pairs emitted before de-duplication */
    long long ccd_pairs;        /* distinct candidate pairs through the narrowphase */
    long long ccd_contacts;     /* contacts found */
    long long ccd_resolved;     /* contacts x passes evaluated by the response */
    long long ccd_ee_raw_pairs; /* edge-edge candidates before de-duplication */
    long long ccd_ee_pairs;     /* distinct edge-edge candidates */
    long long ccd_ee_contacts;  /* edge-edge contacts found */
    /* bytes the LDS-tile launches must move as designed: per launched tile its
       vertex list and loaded state (16 B per vertex, 32 B when it advects x and
       v, +16 B with a kept prediction), the prim records of every phase it
       runs (8 B per distance prim, 12 B per bending prim), its owned prims for
       the residual (8 B each, under a tolerance only) and its stores (16 B per
       interior vertex, 64 B when it integrates); counted from the plan for the
       launches that ran */
    double tile_bytes;
    /* waits of the queued tile launch (per_chunk_launches 0 or 2) that gave up
       after their bounded spin instead of seeing what they waited for (a
       neighbour tile's chunk, a cloth's decision, the next item); 0 in a
       correct run (the count is kept since the context was created) */
    long long queue_wait_timeouts;
    /* queued tile launches that made the warm start's keep decision
       themselves (per_chunk_launches 0; counted on the device, graph replays
       included, since the context was created) */
    long long queue_deciding_launches;
} pbd_profile;
PBD_API int pbd_profile_enable(pbd_ctx* ctx, int on);
PBD_API int pbd_profile_read(pbd_ctx* ctx, pbd_profile* out, int reset);
/* Autotuner (no counterpart in the reference).  From the current state, times
 * `trial_substeps` pbd_step substeps (after one untimed substep) for every
 * candidate of block_threads {256, 512, 1024} x tile_sweeps {1, 2, 4, 5
 * that are <= iterations} x tile_max_verts {device maximum, 4096, 2048}
 * and, with collisions enabled, the hash table size (table_bits - 2 ..
 * table_bits + 2 of the hashed grid), restoring x and v after every trial.
 * The fastest candidate becomes the solver configuration and is returned in
 * `best` (and `best_table_bits`, -1 without collisions).  No candidate
 * changes results: tile plans keep the colour order, hashed candidate sets
 * do not depend on the table size, and with a residual tolerance only
 * tile_sweeps == check_every is tried. */
PBD_API int pbd_autotune(pbd_ctx* ctx, int trial_substeps, pbd_solver_params* best, int* best_table_bits,
                         double* best_ms_per_substep, int* trials);
/* Plain device copy of `bytes` (read once, written once), `reps` times; the
 * HBM calibration kernel for achievable bandwidth and counter scaling. */
PBD_API int pbd_bench_copy(pbd_ctx* ctx, long long bytes, int reps, double* ms_per_copy);

#ifdef __cplusplus
}
#endif
#endif
