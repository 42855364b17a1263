/*
 * eolc.h -- C ABI of the MI355X (gfx950) per-step cloth hot path.
 *
 * Replaces, for solver "none", the per-step work EOL-Cloth does in
 *   Forces::fill            Forces.h:26-45, Forces.cpp:912-930
 *   Cloth::solve            Cloth.cpp:343-357 (b = -(M v + h f))
 *   GeneralizedSolver::velocitySolve   GeneralizedSolver.h:26-30, .cpp:107-172
 *   Cloth::step update loop Cloth.cpp:394-410 (v = solution, x += h v)
 *   CD / CD2                Collisions.h:9-11, Collisions.cpp:11-77
 *   btc::pointTriCollision  boxTriCollision.h:151-158
 *   btc::boxTriCollision    boxTriCollision.h:123-147
 * Plain pointers and sizes only; every call returns an eolc_status and never
 * aborts.  All arrays are caller-owned host memory unless a function says
 * otherwise; the context owns its device buffers.  One context per host
 * thread; calls are synchronous.
 */
#ifndef EOLC_H
#define EOLC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct eolc_ctx eolc_ctx;

typedef enum {
    EOLC_OK = 0,
    EOLC_ERR_INVALID = 1,     /* bad argument, shape or call order */
    EOLC_ERR_HIP = 2,         /* HIP runtime error */
    EOLC_ERR_NOT_SPD = 3,     /* PCG met p^T A p <= 0 (indefinite M + h^2 K) */
    EOLC_ERR_MAXIT = 4,       /* iteration cap reached before the tolerance */
    EOLC_ERR_OVERFLOW = 5,    /* caller buffer too small; required size returned */
    EOLC_ERR_NONFINITE = 6,   /* NaN/Inf met in the state or the solve */
    EOLC_ERR_DEGENERATE = 7,  /* non-positive material-space area (CW face) */
    EOLC_ERR_NO_DEVICE = 8,
    EOLC_ERR_NO_MEMORY = 9    /* a host allocation failed (nothing is thrown across the ABI) */
} eolc_status;

/* option flags (eolc_set_options) */
#define EOLC_OPT_INCLUDE_BENDING_FORCE 0x1 /* add f_bend to f; the reference's non-EOL
                                              branch drops it (Forces.cpp:885-908) */
#define EOLC_OPT_SPD_PROJECT_BENDING   0x2 /* clamp each hinge Hessian's negative
                                              eigenvalues to 0 before assembly */

/* solve modes (eolc_velocity_solve / eolc_step) */
#define EOLC_SOLVE_FP64_BJ       0  /* FP64 3x3 block-Jacobi PCG */
#define EOLC_SOLVE_MIXED_BJ      1  /* FP32 matrix/vectors, FP64 dots and true residual,
                                       reliable updates; block-Jacobi.  When the FP32
                                       iteration stalls (the FP32 operator no longer
                                       resolves the system) the solve continues in FP64
                                       from its iterate (restarts from its initial guess
                                       after linear coarse fields, eolc_set_ml_linear),
                                       and the context's next mixed solves start in FP64
                                       (eolc_set_mixed_stall) */
#define EOLC_SOLVE_MIXED_LEARNED 2  /* as MIXED_BJ with the learned blocks P_i (weights
                                       from eolc_learned_set_weights); falls back to
                                       block-Jacobi after twice the iterations of the last
                                       converged block-Jacobi solve on the context, unless
                                       the network's output layer is zero (its blocks are
                                       then the block-Jacobi ones) */
#define EOLC_SOLVE_FP64_JACOBI   3  /* FP64 scalar-Jacobi PCG (Eigen's default) */
#define EOLC_SOLVE_MULTILEVEL 0x10  /* flag, OR-ed into any mode: the mode's fine blocks plus
                                       the coarse levels of an additive multilevel
                                       preconditioner over an aggregation hierarchy of the
                                       vertex graph (built on first use per topology); the
                                       coarse terms carry the smooth modes that limit the
                                       block preconditioners.  MIXED_LEARNED's fallback
                                       keeps the flag. */

/* Reliable-update trigger of the mixed modes (default 0.5; 0.2 for solves with
 * EOLC_SOLVE_MULTILEVEL, whose cycles are shorter, and 0.1 for the batched solve,
 * whose FP64 residuals run inside the persistent launch): the FP32 correction is folded
 * into the FP64 solution and the residual replaced by the FP64 true residual
 * whenever the recursive residual has dropped by this factor since the last update.
 * A value set here applies to every mode. */
int eolc_set_mixed_delta(eolc_ctx *ctx, double delta);
/* Stiff-block threshold of the FP32 patch records (SpMV variants 25-28), relative to
 * dampingB h^2 E (default 10; 0 turns it off).  A block whose largest entry exceeds
 * it is applied from its FP64 values by both of its rows, because rounded to FP32 a
 * block of a collapsed element (up to 1e10 times the others) perturbs the soft
 * directions of the matrix by more than their own stiffness.  No counterpart in the
 * reference (its solve is FP64 throughout). */
int eolc_set_stiff_threshold(eolc_ctx *ctx, double rel);

/* Stall limits of the mixed modes (no counterpart in the reference).  The FP32
 * iteration is taken to have stalled when a reliable-update cycle runs longer than
 * max(min_cycle, 10x the mean completed cycle) -- first_cycle for the first cycle --
 * or when three reliable updates in a row each reduce the true residual by less than
 * half the cycle's target.  The solve then continues in FP64 (see
 * EOLC_SOLVE_MIXED_BJ), and the next `sticky` mixed solves on the context start in
 * FP64 (a state upload ends that).  Defaults 10000, 1000, 8. */
int eolc_set_mixed_stall(eolc_ctx *ctx, int first_cycle, int min_cycle, int sticky);

/* Linear coarse fields of EOLC_SOLVE_MULTILEVEL (no counterpart in the reference):
 * on the levels >= 3 each aggregate carries, per component, the basis (1, (X - c)/s)
 * of the material coordinates instead of the constant (9 x 9 Galerkin blocks), so a
 * coarse correction is linear on the aggregate and resolves the sheet's smooth
 * bending modes better: 20-25% fewer iterations at 512^2 and 1024^2 on most states.
 * `where`: 0 never, 1 (default) in FP32 iterations only -- FP64 solves, the FP64
 * restart of a stalled mixed solve and its sticky successors keep the translations,
 * because on states with collapsed elements an FP64 solve with the fields took up to
 * 3x the iterations of the translations alone -- 2 in every mode.  With the stiff
 * blocks in FP64 (eolc_set_stiff_threshold) they keep the mixed solves of the
 * benchmark trajectory in FP32 where the translations stall.  They apply when the
 * levels >= 3 hold at most 512 aggregates (a grid of about 1.4 M vertices); larger
 * hierarchies and the batched solve keep the translations. */
int eolc_set_ml_linear(eolc_ctx *ctx, int where);

/* Velocity transfer after a remesh, non-EOL vertices: Cloth::velocityTransfer
 * (Cloth.cpp:195-230, 265-281).  Each new vertex takes the velocity of the first old
 * vertex (index order) within material-space distance 1e-6 (src = its index), else
 * the barycentric combination over the first old face that ArcSim's is_inside
 * accepts (face = its index, bary its coordinates), found through uniform grids
 * instead of the reference's full scans; results are bitwise the scans'.  u: 2
 * material coordinates per vertex, v: 3 per vertex, tri_old: 3 per face; bary (3 per
 * new vertex) and ms (device time) may be NULL.  A new vertex with neither gets
 * face -1, velocity 0 and EOLC_ERR_INVALID (the reference dereferences a null face). */
int eolc_velocity_transfer(eolc_ctx *ctx, int n_old, const double *u_old, const double *v_old, int f_old,
                           const int *tri_old, int n_new, const double *u_new, double *v_new, int *src, int *face,
                           double *bary, double *ms);

/* The same with EOL nodes in the old mesh (Cloth.cpp:254-262): a new vertex placed
 * by barycentric averaging takes, from each vertex j of its old face that was an
 * EOL node (eol_old[j] = its EoL_index, else -1), the velocity v_j - F V_j instead
 * of v_j, F the old face's deformation gradient (deform_grad, UtilEOL.cpp:13-28)
 * at the old world positions x_old (3 per vertex) and V_j = V_old[2 eol_old[j] ..]
 * the node's material-space velocity.  Vertices matched to an old vertex still
 * take its v (the caller copies the EOL part with src, as the reference copies
 * vert->v).  The other EOL branches (WasEOL, NewEOLFromSplit, a new EOL node's
 * split) need the remesher's node states and stay with it. */
int eolc_velocity_transfer_eol(eolc_ctx *ctx, int n_old, const double *u_old, const double *x_old,
                               const double *v_old, const int *eol_old, int n_eol, const double *V_old,
                               int f_old, const int *tri_old, int n_new, const double *u_new, double *v_new,
                               int *src, int *face, double *bary, double *ms);

/* Aggregation hierarchy of EOLC_SOLVE_MULTILEVEL (no counterpart in the reference:
 * a preconditioner of the solver-"none" solve).  eolc_ml_info builds it if needed
 * and reports the number of coarse levels and the aggregates per level: counts[0] =
 * vertices, counts[l] for l = 1..levels (room for EOLC_ML_MAX_LEVELS + 1 ints).
 * eolc_ml_aggregates: the level-l aggregate of every vertex (N ints), l = 1..levels. */
#define EOLC_ML_MAX_LEVELS 8
int eolc_ml_info(eolc_ctx *ctx, int *levels, int *counts);
int eolc_ml_aggregates(eolc_ctx *ctx, int level, int *agg);
/* The same hierarchy without a context or a GPU, for a block pattern given as vertex
 * CSR (rowptr: N + 1 ints, col: the stored blocks' columns; the diagonal may be
 * present).  levels, counts as above; agg (may be null): room for
 * EOLC_ML_MAX_LEVELS x N ints, level l's map at (l - 1) N. */
int eolc_ml_partition(int N, const int *rowptr, const int *col, int *levels, int *counts, int *agg);

/* z = P r with the preconditioner of `mode` (any solve mode, with or without
 * EOLC_SOLVE_MULTILEVEL) for the matrix of the last eolc_forces_fill, fixed DOFs
 * handled as in eolc_velocity_solve: their entries of r are ignored and their
 * entries of z are 0.  r, z: 3N doubles (mixed modes apply P in FP32).  For tests
 * and for callers that run their own Krylov method. */
int eolc_precond_apply(eolc_ctx *ctx, int n_fixed, const int *fixed_dof, int mode, const double *r, double *z);

/* matrices (eolc_download_csr) */
#define EOLC_MAT_M   0
#define EOLC_MAT_MDK 1

typedef struct {
    int iters;              /* inner PCG iterations */
    int refinements;        /* FP64 true-residual replacements (mixed modes) */
    int status;             /* eolc_status of the solve */
    int used_fallback;      /* bit 0: the learned mode fell back to block-Jacobi; bit 1:
                               the mixed solve ran (part of) its iterations in FP64
                               after its FP32 iteration stalled, or started in FP64
                               because a recent mixed solve on the context had stalled */
    double relres;          /* recursive residual estimate at exit */
    double true_relres;     /* ||b - A v||_2 / ||b||_2, FP64, reduced system */
    double ms_assembly;     /* element evaluation + assembly (device time) */
    double ms_rhs;
    double ms_solve;
    double ms_update;
    int launched;           /* PCG iterations enqueued; those past a stop point exit at once */
    int host_syncs;         /* host waits on the solver's stream during the solve */
} eolc_solve_stats;

/* Collision record: btc::Collision (boxTriCollision.h:49-110) flattened. */
typedef struct {
    double dist;
    double nor1[3], nor2[3], pos1[3], pos2[3], pos1_[3];
    int count1, count2;
    int verts1[3], verts2[3];
    double weights1[3], weights2[3];
    int tri1, tri2;
    int nedge1;             /* valid entries of edge1 */
    int edge1[3];
    int edge2;
    double edgeDir[3];
} eolc_collision;

/* ---- lifetime ------------------------------------------------------------ */
int  eolc_create(int device, eolc_ctx **ctx);
void eolc_destroy(eolc_ctx *ctx);
const char *eolc_last_error(const eolc_ctx *ctx);

/* ---- set-up (once per topology; again after every remesh) ------------------
 * X: material coordinates Vert::u (2N); tri: Face::v[3] node indices (3F,
 * counter-clockwise in X); hinge: interior edges as (n[0], n[1], apex of
 * adjf[0], apex of adjf[1]) (4H), Forces.cpp:692-697.  Builds the 3x3 BSR
 * pattern and the element-to-block gather maps on the host. */
int eolc_set_topology(eolc_ctx *ctx, int N, const double *X, int F, const int *tri,
                      int H, const int *hinge);
/* Material (Cloth.h:28-35): density, Young's modulus, Poisson ratio, bending
 * stiffness, dampingA (unused, Forces.cpp:397), dampingB (scales h^2 K). */
int eolc_set_material(eolc_ctx *ctx, double density, double e, double nu, double beta,
                      double dampingA, double dampingB);
int eolc_set_options(eolc_ctx *ctx, int flags);

/* ---- state ------------------------------------------------------------------ */
int eolc_upload_state(eolc_ctx *ctx, const double *x, const double *v);      /* 3N each */
int eolc_download_state(eolc_ctx *ctx, double *x, double *v);

/* ---- Forces::fill ------------------------------------------------------------
 * Evaluates every element at the device-resident x and assembles f (3N; 3N + 2E
 * with E EOL nodes, see eolc_set_eol), M and MDK = M + dampingB h^2 K on the
 * device. */
int eolc_forces_fill(eolc_ctx *ctx, const double grav[3], double h);
int eolc_download_forces(eolc_ctx *ctx, double *f);
/* ---- EOL nodes (Forces.cpp:177-329, 399-497, 580-683, 746-884) ---------------
 * nodes[k] becomes an EOL node with EoL_index k: from the next forces_fill on,
 * f, M and MDK have 3N + 2n rows, the two material-space DOFs of node nodes[k] at
 * 3N + 2k.  Faces and hinges with an EOL vertex add, through the deformation
 * gradient F of the face (deform_grad, UtilEOL.cpp:13-28; for a hinge the mean
 * over its two faces), f_X = -F^T f_v and the blocks F^T B F, -F^T B, -B F of
 * fillXMI / fillXXMI / fillXxMI / fillxXMI and their bending twins; an EOL hinge
 * also adds its bending force to the x rows (edgeBasedF's EOL branch).  The x-x
 * blocks do not change.  The EOL system goes to the host QP with its EOL
 * constraint rows: velocity_solve and step report EOLC_ERR_INVALID while EOL
 * nodes are set.  n = 0 clears them; set_topology clears them too. */
int eolc_set_eol(eolc_ctx *ctx, int n, const int *nodes);

/* Scalar CSR (3N x 3N, sorted columns; 3N + 2E with E EOL nodes) of M or MDK.
 * On EOLC_ERR_OVERFLOW, *nnz holds the required capacity. */
int eolc_download_csr(eolc_ctx *ctx, int which, int64_t *rowptr, int *col, double *val,
                      int64_t cap, int64_t *nnz);

/* ---- GeneralizedSolver::velocitySolve for solver "none" ----------------------
 * Solves MDK v = M v_old + h f on the DOFs not listed in fixed_dof, with
 * v[fixed_dof[i]] = fixed_val[i] (the QP's single-DOF equality rows,
 * Constraints.cpp:114-119; at most 3N entries).  v_old is the device-resident
 * velocity; the solution replaces it.  Warm start from v_old.
 * EOLC_OK means the FP64 true residual of the returned v meets the tolerance:
 * ||b - MDK v|| <= tol ||b|| on the free DOFs, stats->true_relres.  In every mode
 * the recursive residual only triggers that check; when the check fails the true
 * residual replaces the recursive one and the iteration continues (each such
 * replacement counts in stats->refinements).  On any other status v holds the last
 * iterate (with its fixed values), and stats->true_relres is its residual. */
int eolc_velocity_solve(eolc_ctx *ctx, int n_fixed, const int *fixed_dof,
                        const double *fixed_val, double tol, int maxit, int mode,
                        eolc_solve_stats *stats);

/* ---- Cloth::step (Cloth.cpp:359-413, no remesh, no EOL) ----------------------
 * forces_fill, rhs, velocity solve with fixed nodes
 * (v_fixed = 0.99 v_old + scripted velocity, formed on the device from the
 * resident v_old), then x += h v.  When the solve fails (any status but EOLC_OK)
 * the resident x and v are left as they were before the call; the stats describe
 * the failed solve. */
int eolc_step(eolc_ctx *ctx, const double grav[3], double h, int n_fixed_nodes,
              const int *fixed_nodes, const double *scripted_vel /* 3 per node or NULL */,
              double tol, int maxit, int mode, eolc_solve_stats *stats);

/* The same step with single-DOF fixed rows, as Constraints::fill adds them per
 * axis from a FixedList (Constraints.cpp:472-496: corner k, axis a fixed when
 * c_k(a) == 1, scripted velocity c_k(a + 3)): v[fixed_dofs[i]] =
 * 0.99 v_old[fixed_dofs[i]] + scripted_vel[i] (scripted_vel: one per DOF or NULL).
 * The fixed values are formed on the device from the resident v_old. */
int eolc_step_dofs(eolc_ctx *ctx, const double grav[3], double h, int n_fixed_dofs,
                   const int *fixed_dofs, const double *scripted_vel, double tol, int maxit,
                   int mode, eolc_solve_stats *stats);

/* ---- obstacle motion (Obstacles::step) ---------------------------------------
 * Host-only: ctx-free, no device work.  Frames are 4x4 homogeneous transforms,
 * column-major as Eigen stores Matrix4d.
 * eolc_rigid_integrate: Rigid::integrate (Rigid.cpp:79-112), E1 = E0 exp(h phi^)
 *   with phi = (w, v) a body-frame twist (Rodrigues closed form; a pure
 *   translation h v when |w| <= 1e-10).  E1 may alias E0.
 * eolc_box_step: Box::step (Box.cpp:103-114), the box frame after one step of
 *   length h with velocity vel = (angular, linear) as Box::v holds it: phi is
 *   vel mapped by blockdiag(R^T, R^T), R the frame's rotation.  E_new may alias E.
 * Both return EOLC_ERR_INVALID for null or non-finite arguments. */
int eolc_rigid_integrate(const double E0[16], const double phi[6], double h, double E1[16]);
int eolc_box_step(const double E[16], const double vel[6], double h, double E_new[16]);

/* ---- CD2 / CD ------------------------------------------------------------------
 * Proximity queries of the device-resident cloth positions against P points
 * (pts, nrm: 3P) and B boxes (whd: 3B; E: 16B, each a column-major 4x4 frame as
 * Eigen stores Matrix4d).  eol_points = 0 gives CD2 (Collisions.cpp:55-77),
 * 1 gives CD (Collisions.cpp:11-53, with its box feature-index remap; box_npts
 * and box_nedges are Box::num_points / num_edges).  Records are written in the
 * reference's order.  On EOLC_ERR_OVERFLOW *count is the required capacity. */
int eolc_collide(eolc_ctx *ctx, double threshold, int P, const double *pts, const double *nrm,
                 int B, const double *whd, const double *E, int eol_points, int box_npts,
                 int box_nedges, eolc_collision *out, int cap, int *count);

/* The same query with no device allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (tight per-step loops, harnesses
 * that forbid hipMalloc).  Every device buffer of the call is carved from a
 * file-scope device array of EOLC_COLLIDE_SCRATCH_MIB MiB per GPU, held from the
 * loading of the library until the process ends; eolc_collide allocates and frees
 * its buffers on every call instead.  A query that needs more (about 0.6 KiB per
 * cloth vertex; meshes above roughly 350,000 vertices) returns EOLC_ERR_INVALID
 * with the reason, before any device work when the size alone decides it.  The
 * array is shared by every context on a GPU: calls on one GPU from several host
 * threads of a process are serialised by a lock; other processes have their own
 * copy.  Same arguments, records and errors as eolc_collide. */
#define EOLC_COLLIDE_SCRATCH_MIB 256
int eolc_collide_noalloc(eolc_ctx *ctx, double threshold, int P, const double *pts, const double *nrm,
                         int B, const double *whd, const double *E, int eol_points, int box_npts,
                         int box_nedges, eolc_collision *out, int cap, int *count);

/* ---- Constraints::fill, collision rows (Constraints.cpp:417-470) --------------
 * Inequality rows A_ineq v >= 0 of the QP from CD2 records, in record order:
 * (count1, count2) = (3, 1): -nor1 on the DOFs of verts2[0]; (2, 2): -nor2 *
 * weights2[j] on verts2[j], j < 2; (1, 3): -nor2 * weights2[j] on verts2[j], j < 3;
 * other records give no row; rows touching an EOL node (eol[v] != 0; eol may be
 * NULL) are skipped.  Scalar CSR over 3N columns, columns increasing within a row,
 * duplicates summed (what setFromTriplets builds); b_ineq is zero.  Host-only: ctx
 * may be NULL (it only receives the error message).  On EOLC_ERR_OVERFLOW, *rows
 * and *nnz are the required sizes; pass NULL arrays and zero capacities to query. */
int eolc_collision_constraints(eolc_ctx *ctx, const eolc_collision *recs, int n, int N,
                               const unsigned char *eol, int64_t *rowptr, int *col, double *val,
                               int cap_rows, int64_t cap_nnz, int *rows, int64_t *nnz);

/* ---- learned block preconditioner ----------------------------------------------
 * No counterpart in the reference (its solver "none" aborts; Eigen's CG there uses
 * a scalar-Jacobi preconditioner).  A message-passing network over the vertex
 * graph of the current MDK emits per-vertex corrections T_i (lower-triangular,
 * 0.5 tanh of 6 outputs) of the block-Jacobi factor:
 *     P_i = G_i G_i^T,  G_i = chol(D_i^-1) (I + T_i)
 * (SPD for every output; block-Jacobi at output 0).  Parameters: a flat float
 * vector of EOLC_LEARNED_NPARAMS in the order W_in[16][64], b_in[64],
 * {W_l[128][64], b_l[64]} for l = 0..2, W_out[64][6], b_out[6].
 * All calls below use the MDK of the last forces_fill / step and treat the DOFs in
 * fixed_dof as removed (identity rows/columns, as the solve does). */
#define EOLC_LEARNED_NPARAMS 26246

/* Set / read the parameters.  Setting them resets the Adam state. */
int eolc_learned_set_weights(eolc_ctx *ctx, const float *theta, int n);
int eolc_learned_get_weights(eolc_ctx *ctx, float *theta, int n);

/* Network outputs o (6 per vertex, may be NULL) and blocks P (9 per vertex,
 * row-major, may be NULL) for the current system. */
int eolc_learned_infer(eolc_ctx *ctx, int n_fixed, const int *fixed_dof, float *out, float *P);

/* Training loss 1 - (z.r)^2 / ((z.A z)(r.e)), r = A e, z = P r, for an error
 * vector e (3N, fixed entries ignored) and its gradient (EOLC_LEARNED_NPARAMS
 * floats, may be NULL) with respect to the parameters. */
int eolc_learned_loss_grad(eolc_ctx *ctx, int n_fixed, const int *fixed_dof, const double *e,
                           double *loss, float *grad);

/* on != 0: the training loss above uses z = P r + the coarse corrections of
 * EOLC_SOLVE_MULTILEVEL for r, the preconditioner the blocks are used in by
 * EOLC_SOLVE_MIXED_LEARNED | EOLC_SOLVE_MULTILEVEL (the gradient flows through P r;
 * the coarse part does not depend on the parameters).  Default off. */
int eolc_learned_set_multilevel(eolc_ctx *ctx, int on);

/* `steps` Adam steps (lr, beta1 0.9, beta2 0.999) on the current system and e;
 * losses[k] (may be NULL) is the loss before step k.  The Adam state persists
 * across calls until the weights are set again. */
int eolc_learned_train(eolc_ctx *ctx, int n_fixed, const int *fixed_dof, const double *e, int steps,
                       double lr, double *losses);

/* ---- batched cloths (no counterpart in the reference) ---------------------------
 * B independent cloths in one context, solved together.  Instance b has Nb[b]
 * vertices, Fb[b] faces, Hb[b] hinges and nfixb[b] fixed nodes (all DOFs fixed,
 * v = 0.99 v_old + scripted velocity as in eolc_step, the scripted part set by
 * eolc_set_batch_scripted); X, tri, hinge and fixed are the instances' arrays one
 * after the other with instance-local vertex indices.  Each instance is padded
 * to a multiple of 64 vertices with isolated fixed vertices: state arrays
 * (eolc_upload_state / eolc_download_state) use the padded numbering, instance b
 * at vertices [vert_off[b], vert_off[b+1]) (vert_off: B+1 outputs).
 * Replaces eolc_set_topology; density, Poisson ratio and damping come from
 * eolc_set_material, Young's modulus and bending stiffness per instance from
 * eolc_set_batch_material. */
int eolc_set_batch_topology(eolc_ctx *ctx, int B, const int *Nb, const double *X, const int *Fb,
                            const int *tri, const int *Hb, const int *hinge, const int *nfixb,
                            const int *fixed, int *vert_off);
int eolc_set_batch_material(eolc_ctx *ctx, const double *E, const double *beta);

/* One step of every instance (forces, rhs, velocity solve to tol per instance,
 * x += h v).  The solve is one persistent launch: one workgroup per instance runs
 * that instance's mixed-precision PCG (FP32 SpMV and block-Jacobi, FP64 dots,
 * FP64 true residual at reliable updates) and leaves when the instance has
 * converged.  The whole step is captured in a HIP graph on first use and replayed
 * (re-captured when h, tol, maxit, gravity or the launch variant change) unless
 * flags has EOLC_BATCH_EAGER.  stats (B entries, may be NULL) per instance; the
 * return value is the first non-OK instance status. */
#define EOLC_BATCH_EAGER 0x1
int eolc_batch_step(eolc_ctx *ctx, const double grav[3], double h, double tol, int maxit, int flags,
                    eolc_solve_stats *stats);

/* Scripted velocities of the batch's fixed nodes (Cloth::updateFix / FixedList,
 * Constraints.cpp:114-119): vel holds 3 values per node of the `fixed` arrays of
 * eolc_set_batch_topology, in that order (NULL: all zero, the default).  Later
 * steps fix v = 0.99 v_old + vel on those nodes.  The values live in a device
 * buffer that the captured step reads on every replay: changing them at a
 * schedule switch costs one small copy and no re-capture. */
int eolc_set_batch_scripted(eolc_ctx *ctx, const double *vel);

/* Multilevel coarse levels (EOLC_SOLVE_MULTILEVEL's hierarchy, which never spans two
 * cloths) inside the batched persistent solve: on != 0 adds them to every instance's
 * block-Jacobi preconditioner; applies to later eolc_batch_step / _autotune calls. */
int eolc_set_batch_multilevel(eolc_ctx *ctx, int on);

/* Autotuner of the batched solve: times the persistent solve with workgroup sizes
 * 256, 512 and 1024 (ms: 3 entries, may be NULL) on the current state's system
 * and keeps the fastest.  The state is left unchanged.  Results do not depend on
 * the workgroup size (reductions are per 64-row slice, then in slice order). */
#define EOLC_BATCH_VARIANTS 3
int eolc_batch_autotune(eolc_ctx *ctx, const double grav[3], double h, double tol, int maxit, int *best,
                        double *ms);
int eolc_set_batch_variant(eolc_ctx *ctx, int variant);

/* ---- measurement -------------------------------------------------------------
 * Average device time (HIP events on the context's stream) of one launch of the
 * PCG SpMV kernel (precision 64 or 32) over `reps` back-to-back launches on the
 * current matrix.  For roofline reporting; leaves the solver state undefined. */
int eolc_time_spmv(eolc_ctx *ctx, int precision, int reps, double *avg_ms);
/* Sizes of the patch layout of the current topology, for the algorithmic bytes of
 * the patch SpMV variants (no counterpart in the reference): info[0] = 1 when the
 * layout applies, [1] records read per product (upper blocks plus lower blocks whose
 * columns lie in another patch), [2] lower entries (2 bytes each), [3] storage
 * positions (8 bytes of row index and counts each), [4] the cross blocks among [1],
 * [5] LDS slots of the largest patch, [6] window vertices over all patches (each
 * read once per product with its 4-byte index), [7] the largest window.  n: entries
 * of info (at most 8 are written). */
int eolc_spmv_layout_info(eolc_ctx *ctx, int64_t *info, int n);

/* Average device time of one launch of a streaming copy kernel (16-byte loads and
 * stores) reading and writing `bytes` each, over `reps` launches: the known-size
 * kernel for counter calibration and the achievable-bandwidth reference. */
int eolc_stream_copy(eolc_ctx *ctx, int64_t bytes, int reps, double *avg_ms);
/* The same for a streaming read of `bytes` (4-byte non-temporal loads): the
 * achievable read bandwidth, the reference for read-dominated kernels. */
int eolc_stream_read(eolc_ctx *ctx, int64_t bytes, int reps, double *avg_ms);

/* ---- launch autotuner (no counterpart in the reference) -----------------------
 * Times each SpMV launch variant (entries per loop trip, column-index prefetch,
 * non-temporal matrix loads, waves per SIMD, and the symmetric layout that stores
 * each off-diagonal block once) for `reps` launches on the current matrix and
 * keeps the fastest for later solves of that precision (32 or 64) on this context.
 * Variants 19-24 stage each workgroup's neighbour values in LDS; they apply when
 * every workgroup's rows and their columns span at most 5,440 consecutive vertices
 * (a locally numbered mesh, e.g. a row-major grid up to about 1,100 vertices wide);
 * otherwise the autotuner reports them as infinite and eolc_set_spmv_variant
 * refuses them.  Variants 25-28 read the patch layout: the vertices are cut into
 * one compact patch per workgroup (coordinate bisection of X), and a block whose
 * two vertices share a patch is read once, its transposed product passed to the
 * other row through LDS; they apply when every patch needs at most 6,750 LDS entries
 * (the same reporting and refusal otherwise).  26 and 28 load each row's records
 * two ahead instead of one (8 and 4 waves per SIMD).  All variants give bitwise-identical
 * results within a precision, except that variants 18, 22 and 25-28 read the
 * symmetric or patch records in FP64 too and there apply the operator in the
 * difference form of the FP32 product (the same matrix, rounded differently; the
 * true residual is always formed from the assembled FP64 matrix), that the patch
 * variants sum the dot product p.q per patch (other rounding of alpha; their
 * products q are the same bits, except for blocks above the stiff threshold, see
 * eolc_set_stiff_threshold).  ms (may be NULL)
 * receives EOLC_SPMV_VARIANTS average times; *best (may be NULL) the choice.
 * eolc_set_spmv_variant sets the choice directly. */
#define EOLC_SPMV_VARIANTS 29
int eolc_autotune(eolc_ctx *ctx, int precision, int reps, int *best, double *ms);
int eolc_set_spmv_variant(eolc_ctx *ctx, int precision, int variant);

#ifdef __cplusplus
}
#endif
#endif
