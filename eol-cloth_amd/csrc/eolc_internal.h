// Internal declarations shared by the host code and the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>

#include "../../include/eolc.h"

namespace eolc {

// Rows of the 3x3 BSR matrix handled by one workgroup pass of the PCG kernels:
// 16 lanes per block row, 256 threads per workgroup.
constexpr int kPcgWG = 1024;        // threads per workgroup of the PCG row kernels
constexpr int kPcgGrid = 512;       // maximum workgroups of the PCG row kernels
constexpr int kMaxPartials = 2048;  // stride of one partial-sum array

// Device-side solver scalars.  Written only by the last workgroup of a
// reduction kernel (or by a one-thread set-up kernel); read by every
// workgroup of the next kernel.
struct SolverScalars {
    double rz;        // r.z of the current iterate
    double rz_prev;   // r.z of the previous iterate
    double alpha, beta;
    double pq;
    double rr;        // ||r||^2 of the (inner) residual
    double bnorm2;    // ||b_red||^2
    double tol2;      // squared relative tolerance
    double rmax2;     // squared true residual at the last reliable update
    double delta2;    // reliable-update trigger (squared ratio; 0 in FP64 mode)
    double true_rr;   // ||rhs - A v||^2 of the returned solution
    int done;         // 1: stop iterating
    int reason;       // 1 residual below the running threshold, 2 breakdown, 4 non-finite
    int iters;        // completed iterations
    int first_iter;   // iteration that starts with beta = 0
};

struct Topology {
    int N = 0, F = 0, H = 0;
    int64_t nb = 0;          // 3x3 blocks
    int64_t nbp = 0;         // padded stride of the SoA value arrays
    std::vector<int> rowptr, col, diagpos;
    // gather maps (element slots -> blocks, element forces -> vertices)
    std::vector<int> cptr;   // nb + 1
    std::vector<int> clist;  // (slot << 1) | transposed
    std::vector<int> fptr;   // N + 1
    std::vector<int> flist;  // force-vector slot (3F tri slots, then 4H hinge slots)
    std::vector<int> mptr;   // nb + 1, triangle slots only (mass scalars)
    std::vector<int> mlist;  // slot of a triangle block (6f + p), transpose irrelevant
};

template <class T> struct DArr {
    T *p = nullptr;
    size_t n = 0;
};

}  // namespace eolc

struct eolc_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int lpr = 16;               // lanes per block row in the PCG SpMV (8 or 16)
    std::string err;
    int flags = 0;
    // material
    double rho = 0.1, E = 500.0, nu = 0.3, beta = 1e-5, dampA = 0.0, dampB = 1.0;
    bool have_topo = false, have_mat = false, have_state = false, have_forces = false;
    double h_last = 0.0;
    eolc::Topology topo;

    // ---- device buffers
    double *X = nullptr; int *tri = nullptr; int *hinge = nullptr;
    int *rowptr = nullptr, *col = nullptr, *diagpos = nullptr;
    int *cptr = nullptr, *clist = nullptr, *fptr = nullptr, *flist = nullptr;
    int *mptr = nullptr, *mlist = nullptr;
    double *stageK = nullptr;   // (6F + 10H) x 9
    double *stageF = nullptr;   // (3F + 4H) x 3
    double *stageM = nullptr;   // 6F mass scalars
    double *x = nullptr, *v = nullptr, *f = nullptr, *rhs = nullptr;
    double *A64 = nullptr;      // 9 x nbp (SoA)
    float *A32 = nullptr;       // 9 x nbp (SoA)
    double *Mval = nullptr;     // nb mass scalars (M block = m I3)
    double *D64 = nullptr;      // 9 x N  block-Jacobi inverse (SoA)
    float *D32 = nullptr;
    unsigned char *fixmask = nullptr;  // N, bit k = DOF k fixed
    double *vfix = nullptr;     // 3N fixed values (0 where free)
    // PCG work vectors
    double *b64 = nullptr, *xs64 = nullptr, *r64 = nullptr, *z64 = nullptr, *q64 = nullptr;
    double *p64[2] = {nullptr, nullptr};
    float *r32 = nullptr, *z32 = nullptr, *q32 = nullptr, *xl32 = nullptr;
    float *p32[2] = {nullptr, nullptr};
    double *partials = nullptr;   // 6 x kMaxPartials
    unsigned int *counter = nullptr;
    eolc::SolverScalars *scal = nullptr;
    // host pinned copy of the scalars
    eolc::SolverScalars *hscal = nullptr;
    size_t alloc_bytes = 0;
    // collision data (topology only): createEdges table, unit jitter, host triangles
    int *cedges = nullptr;      // 4 per edge: v0, v1, face0, face1 (-1)
    int n_cedges = 0;
    std::vector<int> h_cedges;
    std::vector<int> topo_tri;
    double *cd_unit = nullptr;  // 3N jitter values in (-1, 1)
    int cd_host_decisions = 0;  // angle tests re-decided on the host (diagnostic)
};

namespace eolc {
// kernels (k_elements.hip, k_assemble.hip, k_pcg.hip)
hipError_t launch_face_eval(eolc_ctx *c, double h, const double *grav);
hipError_t launch_hinge_eval(eolc_ctx *c, double h);
hipError_t launch_assemble(eolc_ctx *c, bool want32);
hipError_t launch_rhs(eolc_ctx *c, double h);
hipError_t launch_update_x(eolc_ctx *c, double h);

int pcg_grid(int N);
hipError_t pcg_setup(eolc_ctx *c, bool mixed, int precond_kind);
hipError_t pcg_init(eolc_ctx *c, bool mixed);
hipError_t pcg_chunk(eolc_ctx *c, bool mixed, int k0, int iters);
hipError_t pcg_reliable_update(eolc_ctx *c);
hipError_t pcg_restart_inner(eolc_ctx *c, int k);
hipError_t pcg_finish(eolc_ctx *c, bool mixed);
hipError_t spmv_true_residual(eolc_ctx *c);
hipError_t scal_set(eolc_ctx *c, double tol2, double delta2, int reset);
hipError_t set_fixed(eolc_ctx *c, int n, const int *d_dofs, const double *d_vals, int set);
hipError_t time_spmv(eolc_ctx *c, bool mixed, int reps, float *ms);
void build_collision_edges(eolc_ctx *c, int F, const int *tri);
void build_jitter(std::vector<double> &u, int N);
}  // namespace eolc
