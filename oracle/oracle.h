/*
 * oracle.h -- CPU FP64 restatement of the EOL-Cloth per-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code; it is the checker for the HIP path (tests/, the smoke check
 * and the CPU-baseline leg of bench.py).
 *
 * Every function cites the reference file:line it restates.  The element
 * functions are written from the closed forms of the energies (derived from the
 * mathematics, not transcribed from the generated code) and are pinned by:
 *   - golden vectors produced by the reference's own ComputeMembrane.cpp,
 *     ComputeInertial.cpp and ComputeBending.cpp (tests/golden/elements.npz,
 *     and oracle/_ref/libref_elements.so when it is built);
 *   - central finite differences of W -> f -> K;
 *   - invariants (symmetry, translation null space, rigid-rotation invariance,
 *     rest-state zeros, total mass).
 * Parity status per function is listed in the header of each source file.
 */
#ifndef EOLC_ORACLE_H
#define EOLC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- material (Cloth.h:28-35, parseParams.cpp:193-201) ---------------- */
typedef struct {
    double density;   /* rho, kg/m^2 */
    double e;         /* Young's modulus, Pa */
    double nu;        /* Poisson's ratio */
    double beta;      /* bending stiffness */
    double dampingA;  /* parsed, never used on the path (Forces.cpp:397) */
    double dampingB;  /* scales h^2 K (Forces.cpp:105) */
} orc_material;

/* ---- mesh (orc_mesh.c) -------------------------------------------------- */
int  orc_grid_mesh(int n, double *X, int *tri);
int  orc_build_hinges(int N, int F, const int *tri, int *hinge /* 4*cap */, int cap);

/* ---- elements (orc_elements.c) ----------------------------------------- */
void orc_inertial(const double *xa, const double *xb, const double *xc,
                  const double *Xa, const double *Xb, const double *Xc,
                  const double *g, double rho, double *W, double *f, double *M);
void orc_membrane(const double *xa, const double *xb, const double *xc,
                  const double *Xa, const double *Xb, const double *Xc,
                  double e, double nu, const double *P, const double *Q,
                  double *W, double *f, double *K);
void orc_bending(const double *x0, const double *x1, const double *x2, const double *x3,
                 const double *X0, const double *X1, const double *X2, const double *X3,
                 double beta, double *W, double *f, double *K);
void orc_face_PQ(const double *xa, const double *xb, const double *xc,
                 const double *Xa, const double *Xb, const double *Xc,
                 double *P /* 2x3 col-major */, double *Q /* 2x2 col-major */);
void orc_poldec(const double *M /* 2x2 col-major */, double *Q /* 2x2 col-major */);

/* ---- assembly (orc_assembly.c) ------------------------------------------ */
typedef struct {
    int n;            /* rows = cols */
    int64_t nnz;
    int64_t *rowptr;  /* n+1 */
    int *col;         /* nnz, sorted within a row */
    double *val;      /* nnz */
} orc_csr;

void orc_csr_free(orc_csr *A);

/* flags */
#define ORC_INCLUDE_BENDING_FORCE 1  /* add f_bend to f (reference drops it: Forces.cpp:885-908) */

int orc_forces_fill(int N, const double *x, const double *X,
                    int F, const int *tri, int H, const int *hinge,
                    const orc_material *mat, const double *grav, double h, int flags,
                    double *f, orc_csr *M, orc_csr *MDK, int nthreads);

/* Forces::fill with EOL nodes (Forces.cpp:177-329, 399-497, 580-683, 746-884):
 * eol_nodes[k] has EoL_index k and two more DOFs at 3N + 2k; f has 3N + 2 n_eol
 * entries and M, MDK that many rows.  Returns 1 for a bad or repeated node. */
int orc_forces_fill_eol(int N, const double *x, const double *X,
                        int F, const int *tri, int H, const int *hinge,
                        const orc_material *mat, const double *grav, double h, int flags,
                        int n_eol, const int *eol_nodes, double *f, orc_csr *M, orc_csr *MDK, int nthreads);

void orc_csr_matvec(const orc_csr *A, const double *x, double *y, int nthreads);

/* ---- solve (orc_solve.c) ------------------------------------------------- */
#define ORC_PC_JACOBI 0
#define ORC_PC_BLOCK_JACOBI 1
typedef struct {
    int iters;
    double relres;       /* recursive residual at exit */
    double true_relres;  /* ||b - A v|| / ||b|| recomputed at exit (reduced system) */
    int status;          /* 0 ok, 1 iteration cap, 2 breakdown (p^T A p <= 0), 3 bad input */
} orc_solve_stats;

int orc_velocity_solve(const orc_csr *MDK, const double *rhs,
                       int nfixed, const int *fixed_dof, const double *fixed_val,
                       const double *v0, double tol, int maxit, int precond,
                       double *v, orc_solve_stats *st, int nthreads);
int orc_dense_velocity_solve(const orc_csr *MDK, const double *rhs,
                             int nfixed, const int *fixed_dof, const double *fixed_val,
                             double *v);

int orc_step(int N, double *x, double *v, const double *X,
             int F, const int *tri, int H, const int *hinge,
             const orc_material *mat, const double *grav, double h, int flags,
             int nfixed_nodes, const int *fixed_nodes, const double *fixed_scripted_vel,
             double tol, int maxit, int precond, orc_solve_stats *st, int nthreads,
             double *phase_seconds /* [5]: elements+assembly, rhs, solve, update, total */);

/* ---- collisions (orc_collide.c) ------------------------------------------ */
typedef struct {
    double dist;
    double nor1[3], nor2[3], pos1[3], pos2[3], pos1_[3];
    int count1, count2;
    int verts1[3], verts2[3];
    double weights1[3], weights2[3];
    int tri1, tri2;
    int nedge1;
    int edge1[3];
    int edge2;
    double edgeDir[3];
} orc_collision;

typedef struct {
    int verts[4];
    int faces[2];
    int internal;
    double angle;
    double normals[6];
} orc_edge;

void orc_mt19937_uniform(uint32_t seed, int count, double lo, double hi, double *out);
/* intersect_triangle3_inc (raytri.cpp:259-316): 1 and *t on a hit */
int  orc_tri3_inc(const double *orig, const double *dir, const double *v0, const double *v1, const double *v2,
                  double *t);
int  orc_create_edges(int N, int F, const int *tri, const double *verts, orc_edge *edges /* cap 3F */);
int  orc_point_tri(double thr, int P, const double *pts, const double *nrm,
                   int N, const double *verts, int F, const int *tri, int eol,
                   orc_collision *out, int cap);
int  orc_box_tri(double thr, const double *whd, const double *E /* 4x4 col-major */,
                 int N, const double *verts, int F, const int *tri, const int *isEOL, int eol,
                 orc_collision *out, int cap);

/* ---- velocity transfer after a remesh (orc_transfer.c) ------------------- */
/* u: 2 per vertex (material space), v: 3 per vertex; per new vertex: src (first old
 * vertex within 1e-6, or -1), face (first enclosing old face, or -1), bary (3).
 * Returns the number of new vertices with neither. */
int orc_velocity_transfer(int n_old, const double *u_old, const double *v_old, int f_old, const int *tri_old,
                          int n_new, const double *u_new, double *v_new, int *src, int *face, double *bary);
/* The same with EOL nodes in the old mesh (Cloth.cpp:254-262): eidx[j] = EoL index
 * of old vertex j or -1, V_old 2 per EOL node, x_old the old world positions. */
int orc_velocity_transfer_eol(int n_old, const double *u_old, const double *x_old, const double *v_old,
                              const int *eidx, const double *V_old, int f_old, const int *tri_old,
                              int n_new, const double *u_new, double *v_new, int *src, int *face, double *bary);

/* ---- input generators (orc_rng.c) --------------------------------------- */
uint64_t orc_splitmix64(uint64_t *state);
void orc_gaussian(uint64_t seed, int64_t count, double sigma, double *out);

#ifdef __cplusplus
}
#endif
#endif
