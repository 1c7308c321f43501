/*
 * pbd_oracle.h -- CPU oracle for the PBD cloth substep.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithms (vivienldy/PositionBasedDynamicsOld, directory
 * PositionBasedDynamicsCUDA/PositionBasedDynamicsCUDA/, abbreviated PBD/ below).
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and only to check results or as the timed CPU baseline.  The product
 * library (libpbd.so) never links or calls it.
 *
 * Every function cites the reference file:line it follows.  Pinning of each
 * function (what independent check fixes it) is listed next to it; functions
 * with no independent check say "parity unpinned".
 *
 * Conventions: positions are AoS float3 (x,y,z) exactly like the reference's
 * BufferVector3f; indices are int32; w[] is the per-vertex weight the
 * reference keeps in massBuffer (1 = free, 0 = pinned).
 */
#ifndef PBD_ORACLE_H
#define PBD_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- mesh / grid (PBD/PBD_cuda.cu:831-869, 19-54, 3-6) ------ */
uint64_t orc_splitmix64(uint64_t* state);
float orc_uniform_pm1(uint64_t* state);
void orc_grid_positions(int rx, int ry, float sx, float sy, float* pos);
void orc_jitter(float* pos, int nv, float amp, uint64_t seed);
int orc_grid_tris(int rx, int ry, int* tris);
int orc_grid_distance_pairs(int rx, int ry, int* pairs);
float orc_distance_powf(const float* a, const float* b);
float orc_distance(const float* a, const float* b);
void orc_rest_lengths(const float* pos, const int* pairs, int ne, float* rest);

/* ---------------- bending (new; Bridson et al. 2003 dihedral) ------------ */
int orc_bending_quads(const int* tris, int nt, int* quads);
float orc_atan2f(float y, float x);
float orc_dihedral(const float* e0, const float* e1, const float* o0, const float* o1);
void orc_dihedral_grad(const float* e0, const float* e1, const float* o0, const float* o1,
                       double* g /* 12: d theta / d (e0,e1,o0,o1) */);
void orc_dihedral_grad_f64(const double* e0, const double* e1, const double* o0,
                           const double* o1, double* theta, double* g);

/* ---------------- adjacency + colouring (PBD/PBD_cuda.cu:155-202, 313-431) */
int orc_prim_adjacency(const int* prim_list /* 2*np (start,count) */, int np,
                       const int* indices, int nv, int* adj_start_num /* 2*np or NULL */,
                       int* adj /* or NULL */);
int orc_color(const int* prim_list, int np, const int* adj_start_num, const int* adj,
              int max_rounds, int* color, int* prd_color);
int orc_color_sort(const int* color, int np, int* sorted_prim, int* worksets /* 2*C */);

/* ---------------- substep (PBD/PBD_cuda.cu:949-1127, 1240-1306, 1453-1573) */
typedef struct {
    float dt;
    float gravity[3];
    float damp_factor;     /* powf(dampingRate, dt), computed once by the caller */
    float wind[3];         /* wind velocity (m/s) */
    float wind_drag;       /* per-second drag coefficient; 0 disables the wind term */
    int sphere_on;
    float sphere_c[3];
    float sphere_r;
    int box_on;
    float box_lo[3];
    float box_hi[3];
} orc_params;

typedef struct {
    int nv;
    float* x;             /* 3*nv positions */
    float* v;             /* 3*nv velocities */
    float* p;             /* 3*nv predicted positions */
    const float* w;       /* nv weights */
    int np;               /* prims (distance: 2 points, bending: 4 points) */
    const int* prim_list; /* 2*np (start,count) */
    const int* indices;
    const float* rest;    /* rest length (distance) or rest angle (bending) */
    const int* sorted;    /* colour-major prim order */
    int ncolors;
    const int* worksets;  /* 2*ncolors (start,count) */
    const float* k_dist;  /* per-sweep stiffness factor for distance prims */
    const float* k_bend;  /* per-sweep stiffness factor for bending prims */
    int k_len;            /* length of the k tables */
} orc_cloth;

void orc_advect(const orc_params* prm, int nv, const float* w, const float* x, float* v, float* p);
void orc_project_distance(const int* ids, float rest, float k, const float* w, float* p);
void orc_project_bending(const int* ids, float rest, float k, const float* w, float* p);
void orc_sweep(orc_cloth* c, int sweep_index);
void orc_colliders(const orc_params* prm, int nv, const float* w, float* p);
void orc_integrate(int nv, float dt, float* x, float* v, const float* p);
void orc_residual(const orc_cloth* c, const float* p, double* max_rel, double* sum_sq, int* count);
void orc64_substeps(const orc_cloth* c, const orc_params* prm, double* x, double* v, double* p, int sweeps,
                    int substeps);
int orc_substep_omp(orc_cloth* c, const orc_params* prm, int cap, int nthreads);
int orc_substep(orc_cloth* c, const orc_params* prm, int cap, int check_every, double tol,
                double* final_residual);
void orc_jacobi_sweep(const orc_cloth* c, int sweep_index, float omega, const float* pin, float* pout);
int orc_substep_jacobi(orc_cloth* c, const orc_params* prm, int cap, float omega, int check_every, double tol,
                       double* final_residual);
void orc_stiffness_table_ref(float s, int n, float* k);
void orc_stiffness_table_const(float s, int iterations, int n, float* k);

/* ---------------- broadphase (PBD/SpatialHashSystem.cpp:88-187, 276-413) -- */
typedef struct {
    int mode;             /* 0 = dense grid (reference), 1 = hashed table */
    float start[3];       /* grid start = centre - size*cell/2 (InitSH) */
    float cell[3];
    int size[3];          /* dense grid dims */
    int table_bits;       /* hashed mode: table = 1 << table_bits */
} orc_grid;

void orc_grid_init(orc_grid* g, const float centre[3], const int size[3], const float cell[3]);
void orc_tri_centroids(const float* pos, const int* tris, int nt, float* cen);
void orc_cell_coords(const orc_grid* g, const float* c, int* gp);
int orc_cell_key(const orc_grid* g, const int* gp);
void orc_tri_keys(const orc_grid* g, const float* pos, const int* tris, int nt, int* keys, int* ids);
void orc_stable_sort_pairs(int* keys, int* vals, int n);
int orc_num_cells(const orc_grid* g);
void orc_cell_ranges(const int* sorted_keys, int nt, int ncells, int* cstart, int* cend);
long long orc_candidate_pairs(const orc_grid* g, const float* pos, const int* tris, int nt,
                              int nv, long long cap, int64_t* out);

/* ---------------- CCD (PBD/CCD.cpp:8-147, 1733-1808, 1996-2090) ----------- */
int orc_solve_quadratic(double a, double b, double c, double x[2]);
double orc_newtons_method(double a, double b, double c, double d, double x0, int init_dir);
int orc_solve_cubic(double a, double b, double c, double d, double x[3]);
void orc_barycentric(const float* p, const float* a, const float* b, const float* c, float* uvw);
int orc_vf_ccd(const float* x0, const float* x1, const float* x2, const float* x3,
               const float* p0, const float* p1, const float* p2, const float* p3,
               float thickness, float* t_out, float* n_out, double* w_out);
int orc_vf_resolve_test(const float* x0, const float* x1, const float* x2, const float* x3,
                        const float* p0, const float* p1, const float* p2, const float* p3,
                        float thickness, float min_n2);
void orc_vf_resolve(const float* x0, const float* x1, const float* x2, const float* x3,
                    float* p0, float* p1, float* p2, float* p3, float thickness);

/* Contact list: candidates that pass the narrowphase, in ascending (v, f)
 * key order.  f indexes the triangle array. */
typedef struct {
    int v, f;
    float t;
    float n[3];
    double w[4];
} orc_contact;

long long orc_swept_pairs(const float* x, const float* p, const int* tris, int nt, int nv, float thickness,
                          int all_pairs, long long cap, int64_t* out);
int orc_vf_cross(const float* x0, const float* x1, const float* x2, const float* x3,
                 const float* p0, const float* p1, const float* p2, const float* p3, float thickness);
long long orc_crossings(const float* x, const float* p, const int* tris, int nt, int nv, float thickness,
                        const float* w, long long cap, int64_t* out);
long long orc_ccd_detect_swept(const float* x, const float* p, const int* tris, int nt, int nv, float thickness,
                               int all_pairs, long long cap, orc_contact* out);
long long orc_ccd_detect(const orc_grid* g, const float* x, const float* p, const int* tris, int nt,
                         int nv, float thickness, long long cap, orc_contact* out);
int orc_edge_table(const int* tris, int nt, int* tedge, int* edges);
void orc_seg_closest(const float* a0, const float* a1, const float* b0, const float* b1, float* s, float* u);
int orc_ee_ccd(const float* x0, const float* x1, const float* x2, const float* x3,
               const float* p0, const float* p1, const float* p2, const float* p3,
               float thickness, float* t_out, float* n_out, double* w_out);
int orc_ee_fix(const float* X0, const float* X1, const float* X2, const float* X3, const float* q0,
               const float* q1, const float* q2, const float* q3, float thickness, float* d);
long long orc_ee_candidates(const orc_grid* g, const float* pos, const int* tris, int nt, const int* tedge,
                            const int* edges, long long cap, int* out);
long long orc_ccd_detect_ee(const orc_grid* g, const float* x, const float* p, const int* tris, int nt,
                            float thickness, long long cap, int* quads, float* t, float* n, double* w);
void orc_ccd_resolve_ee(const orc_contact* c, long long nc, const int* quads, long long nee, const int* tris,
                        const float* x, float* p, const float* w, int nv, float thickness, int passes,
                        float min_n2);
void orc_ccd_resolve(const orc_contact* c, long long nc, const int* tris, const float* x, float* p,
                     const float* w, float thickness, int passes, float min_n2);

/* ---------------- warm-start MLP, FP64 (new) ------------------------------ */
#define ORC_MLP_IN 22
#define ORC_MLP_H 64
#define ORC_MLP_OUT 3
void orc_gelu(double x, double* y, double* dy);
void orc_mlp_forward(const double* W1, const double* b1, const double* W2, const double* b2,
                     const double* W3, const double* b3, const double* x, int n, double* y,
                     double* h1, double* a1, double* h2, double* a2);
double orc_mlp_loss_grad(const double* W1, const double* b1, const double* W2,
                         const double* b2, const double* W3, const double* b3,
                         const double* x, const double* t, int n, double* gW1, double* gb1,
                         double* gW2, double* gb2, double* gW3, double* gb3);
void orc_adam_step(double* w, const double* g, double* m, double* v, int n, double lr,
                   double b1, double b2, double eps, int step);
uint16_t orc_f32_to_bf16(float x);
float orc_bf16_to_f32(uint16_t h);

#ifdef __cplusplus
}
#endif
#endif
