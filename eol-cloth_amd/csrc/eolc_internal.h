// Internal declarations shared by the host code and the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <exception>
#include <new>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/eolc.h"

namespace eolc {

// Rows of the 3x3 BSR matrix handled by one workgroup pass of the PCG kernels:
// 16 lanes per block row, 256 threads per workgroup.
constexpr int kPcgWG = 1024;        // threads per workgroup of the PCG row kernels
constexpr int kPcgGrid = 512;       // maximum workgroups of the PCG row kernels
constexpr int kMaxPartials = 16384; // stride of one partial-sum array (max PCG workgroups)

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

// 3x3 block matrix in sliced ELL form (SELL-C-sigma, C = 64 = one wavefront):
// rows are grouped in slices of 64 consecutive vertices; inside a slice the rows are
// ordered by decreasing length, so at every column position k the rows that still
// have a block form a prefix of the wavefront.  Entry (slice s, position k, lane l)
// lives at slice_off[s] + 64 k + l; storage position 64 s + l holds row srow[..]
// with slen[..] blocks.  The CSR form (rowptr/col, columns sorted) is kept on the
// host for downloads.
constexpr int kSlice = 64;
#ifndef EOLC_SELL_SIGMA
#define EOLC_SELL_SIGMA 64
#endif
constexpr int kSellSigma = EOLC_SELL_SIGMA;   // rows sorted by length inside windows of this many rows
// (64: one slice; a window of 4 slices evens the trip counts of a checkerboard grid
// but spreads each slice's gathers wider and measured 5% slower)

struct Topology {
    int N = 0, F = 0, H = 0;
    int64_t nb = 0;          // 3x3 blocks (CSR)
    std::vector<int> rowptr, col;
    // SELL layout
    int nslices = 0;
    int64_t nbs = 0;         // SELL entries including padding (stride of the SoA value arrays)
    std::vector<int> slice_off, srow, slen, scol, sellpos, sdiag;
    // symmetric FP32 layout (each off-diagonal block once, with its upper row)
    std::vector<int> supper;        // per storage position: number of upper entries
    std::vector<int> uoff, loff;    // per slice: first upper / lower record
    int64_t nup = 0, nlo = 0;       // records including padding
    std::vector<int> e2u;           // per SELL entry: its upper record or -1
    std::vector<int> ucol;          // per upper record: column
    std::vector<int> lrec;          // per lower record: column, upper record of the transposed block
    // gather maps in SELL entry order (element slots -> entries, element forces -> vertices)
    std::vector<int> cptr;   // nbs + 1
    std::vector<int> clist;  // (slot << 1) | transposed
    std::vector<char> has_tri;  // per CSR block: at least one triangle contribution (M pattern)
    std::vector<int> fptr;   // N + 1
    std::vector<int> flist;  // force-vector slot (3F tri slots, then 4H hinge slots)
    // per PCG workgroup (logical index, pcg_grid(N) of them): the vertex range
    // [wwin[2 lw], wwin[2 lw + 1]) that holds its rows and every column they reference;
    // wmax = the widest range (the LDS-staged SpMV variants need wmax <= kStageVerts)
    std::vector<int> wwin;
    int wmax = 0;
    // Patch layout (SpMV layout 5): the vertices are cut into pcg_grid(N) compact
    // patches of at most kPcgWG vertices (recursive coordinate bisection of the
    // material coordinates), one per PCG workgroup, each stored as 16 slices of 64
    // rows at positions 1024 lw + 64 slice + lane.  A row's records are its upper
    // blocks (column > row) in increasing column order, then its lower blocks whose
    // column lies in another patch (cross blocks); a lower block inside the patch is
    // not stored again: its owner (the upper row) leaves the transposed product in
    // LDS for it, and the row's lower list names that LDS slot (or kPatchCross: take
    // the row's next cross record).
    bool pt_ok = false;
    std::string pt_why;                  // why the layout does not apply to this mesh
    std::vector<int> pt_row;             // per storage position: row or -1
    std::vector<int> pt_cnt;             // per storage position: upper records | lower entries << 16
    std::vector<int> pt_roff;            // per slice (+1): first record
    std::vector<int> pt_soff;            // per slice: patch-local LDS slot of its first upper position
    std::vector<int> pt_loff;            // per slice (+1): first lower entry
    std::vector<unsigned short> pt_lo;   // lower entries: LDS slot or kPatchCross
    std::vector<int> pt_e2r;             // per SELL entry: its record or -1
    std::vector<int> pt_e2l;             // per SELL entry with a record: its column's window index
    std::vector<int> pt_woff;            // per patch (+1): first window vertex
    std::vector<int> pt_win;             // windows: the patch's rows in storage order, then the
                                         // other vertices they reference (increasing index)
    int64_t pt_nrec = 0, pt_nlo = 0;
    int pt_slots = 0;                    // LDS slots of the largest patch
    int pt_wmax = 0;                     // largest window
};

// Patch layout limits.  A workgroup's LDS pool holds 3 values per transposed-product
// slot and per window vertex: kPatchPool entries (81,000 B in FP32, two workgroups
// per CU; 162,000 B in FP64, one).  A window holds at most kPatchWin vertices (staged
// through registers before beta is known).  kPatchCross marks a cross lower entry.
constexpr int kPatchPool = 6750;
constexpr int kPatchWin = 2048;
constexpr int kPatchSlots = kPatchPool;   // slot indices stay below this
constexpr unsigned short kPatchCross = 0xffff;

// Preconditioner blocks are symmetric by construction (block-Jacobi inverse by the
// adjugate, learned G G^T) and stored as their 6 upper entries, SoA over vertices:
// D[k * N + i], k = 00, 01, 02, 11, 12, 22.
constexpr int kDsym = 6;
constexpr int kSymIdx[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};   // row-major 3x3 -> stored entry

// z = D r for one vertex, each row as fma(d2, r2, fma(d1, r1, d0 r0)).
template <class T>
__host__ __device__ __forceinline__ void apply_sym3(const T *__restrict__ d, int64_t N, T r0, T r1, T r2, T &z0, T &z1,
                                                    T &z2)
{
    const T d00 = d[0], d01 = d[N], d02 = d[2 * N], d11 = d[3 * N], d12 = d[4 * N], d22 = d[5 * N];
    z0 = fma(d02, r2, fma(d01, r1, d00 * r0));
    z1 = fma(d12, r2, fma(d11, r1, d01 * r0));
    z2 = fma(d22, r2, fma(d12, r1, d02 * r0));
}

template <class T> struct DArr {
    T *p = nullptr;
    size_t n = 0;
};

// Batched cloths: instance b owns vertices [vert_off[b], vert_off[b+1]) (a multiple of
// 64 each, so every SELL slice belongs to one instance; the padding vertices are
// isolated and fixed) and SELL slices [vert_off[b] / 64, vert_off[b+1] / 64).
struct BatchInst {
    int v0, v1;      // vertex range (padded)
    int s0, s1;      // slice range
    int n_real;      // vertices before padding
    int pad_[3];
};
constexpr int kBatchVariants = 3;   // persistent-solve workgroup sizes 256, 512, 1024
// Reliable updates allowed per solve: a residual that has left FP32's range (or
// keeps reporting convergence of the inner iteration without progress) ends the
// solve with EOLC_ERR_MAXIT instead of looping.
constexpr int kMaxRefinements = 1000;
// Mixed-precision stall: a reliable-update cycle longer than max(stall_min, 10x the mean
// completed cycle) (stall_first for the first cycle), or three reliable updates in a row
// that each reduce the true residual by less than half the cycle's target (a factor
// delta), end the FP32 iteration; the solve continues in FP64, and the next
// stall_sticky mixed solves on the context start in FP64 (defaults 10000, 1000, 8;
// eolc_set_mixed_stall).

// EOL nodes (Forces.cpp:177-329, 399-497, 580-683, 746-884): node nodes[k] gets two
// material-space DOFs at 3N + 2k.  The maps below are built by eolc_set_eol on the
// host; k_eol.hip fills the values after each forces_fill.
//   stage:  per EOL element e (faces, then hinges, with an EOL vertex), for each EOL
//           local vertex v (in local order) a record of 4 nE + 6 nv values: the 2x2
//           blocks (X_v, X_w) for the EOL locals w, then the 2x3 blocks (X_v, x_u)
//           for every local u; forces: 2 per EOL local vertex.
//   X rows: CSR over rows 3N .. 3N + 2E - 1 (stored from 0), full columns, each
//           entry's contributions in element order (setFromTriplets' fold order).
struct EolState {
    int E = 0;
    std::vector<int> nodes;              // EoL_index -> node
    int ne = 0;                          // EOL elements
    std::vector<int64_t> rowptr;         // 2E + 1
    std::vector<int> col;                // columns of the X-row entries
    std::vector<char> has_face;          // an entry of the M pattern (a face contributes)
    int *d_elem = nullptr;               // ne: f for a face, F + h for a hinge
    int *d_eidx = nullptr;               // N: EoL_index or -1
    int64_t *d_soff = nullptr, *d_foff = nullptr;   // ne + 1 stage offsets
    int64_t *d_cptr = nullptr;           // entries + 1
    int *d_clist = nullptr;              // stage indices per entry
    int *d_fptr = nullptr, *d_flist = nullptr;      // 2E + 1, force stage indices per X DOF
    int nxv = 0;                         // vertices of EOL hinges (their bending force joins f)
    int *d_xv = nullptr, *d_xvptr = nullptr, *d_xvslot = nullptr;
    double *d_stK = nullptr, *d_stM = nullptr, *d_stF = nullptr;
    double *d_A = nullptr, *d_M = nullptr, *d_fX = nullptr;   // X-row MDK and M values, X forces
};

struct BatchState {
    int B = 0;
    std::vector<BatchInst> inst;
    std::vector<int> order;          // launch order: largest instance first
    std::vector<int> fixed_nodes;    // padded vertex indices with all DOFs fixed (incl. padding)
    std::vector<int> user_fixed;     // the caller's fixed nodes (v = 0.99 v_old + scripted)
    BatchInst *d_inst = nullptr;
    int *d_order = nullptr;
    int *d_user_fixed = nullptr;
    int *d_all_fixed = nullptr;      // every fixed node (corners and padding): the batch's fixmask
    int n_all_fixed = 0;
    double *d_scripted = nullptr;    // 3 per user-fixed node: scripted velocity (zero by default)
    double *slice_part = nullptr;    // per-slice FP64 partial sums (4 x nslices)
    int *stats = nullptr;            // per instance: iters, refinements, status, spare
    double *dstats = nullptr;        // per instance: bnorm2, true_rr, relres, spare
    int variant = 2;                 // workgroup size choice
    bool ml = false;                 // multilevel coarse levels in the persistent solve
    int *d_ml_lo = nullptr;          // (B + 1) x (kMLMaxLevels + 1) aggregate ranges per instance
    bool graph_ml = false;
    hipGraphExec_t graph = nullptr;  // captured step
    double graph_h = 0.0, graph_tol = 0.0;
    double graph_grav[3] = {0, 0, 0};
    int graph_maxit = 0, graph_variant = -1;
};

// Aggregation hierarchy of the additive multilevel preconditioner (k_ml.hip).
// Level l >= 1 partitions the vertices into aggregates, each level-l aggregate a
// union of level-(l-1) aggregates.  Aggregate ids are hierarchical: the members of
// every level-1 aggregate are a contiguous range of `hord` (vertices in hierarchy
// order), and the children of every level-l aggregate (l >= 2) a contiguous range
// of level-(l-1) ids.  Per-aggregate arrays of all levels are concatenated: level l
// starts at aggregate `off[l]`.
constexpr int kMLMaxLevels = 8;
constexpr int kMLLinFrom = 3;     // first level with linear coarse fields
constexpr int kGal9Rows = 2048;   // hord positions per set-up chunk of a 9 x 9 Galerkin block
constexpr int kRestrict3Parts = 4;   // restriction workgroups per level-3 aggregate (linear fields)
constexpr int kCoarse9WG = 512;      // the linear fields' coarse kernel: at most this many aggregates on levels >= 3
struct MLLevels {
    int L;                        // coarse levels built (0: none)
    int n[kMLMaxLevels + 1];      // aggregates per level (n[0] = vertices)
    int off[kMLMaxLevels + 2];    // first aggregate of level l in the concatenated arrays
    int coff[kMLMaxLevels + 1];   // first entry of level l's child ranges in cptr (l >= 2)
};

struct MLState {
    bool built = false;           // hierarchy matches the current topology
    MLLevels lv{};
    int *hord = nullptr;          // N: vertices in hierarchy order
    int *ptr1 = nullptr;          // n_1 + 1: member range of each level-1 aggregate in hord
    int *agg1 = nullptr;          // N: level-1 aggregate of each vertex
    int *gpos = nullptr;          // N: SELL storage position of each row
    int *cptr = nullptr;          // child ranges of levels 2..L (n_l + 1 entries each)
    int *par = nullptr;           // parent (level l+1 id) of each aggregate of levels 1..L-1
    unsigned char *lam = nullptr; // per SELL entry: first level whose aggregate holds both
                                  // vertices (0: diagonal, L+1: none, 255: padding)
    double *V = nullptr;          // set-up sums: level l holds L x 9 x n_l
    double *Dinv = nullptr;       // 6 per aggregate: inverse of the Galerkin diagonal block
    double *R = nullptr;          // 3 per aggregate: restricted residual
    double *C = nullptr;          // 3 per aggregate: accumulated coarse correction; for a
                                  // level-1 aggregate a (L >= 2): D_1^-1 R_1(a) + D_2^-1 R_2(parent)
    int *anc3 = nullptr;          // N (L >= 3): global index of each vertex's level-3 aggregate
    int *kptr2 = nullptr;         // n_2 + 1 (L >= 2): member range in hord of each level-2 aggregate
    // Linear coarse fields on the levels >= 3 (single-cloth solves; the batched solve
    // keeps the translations): per aggregate of every level its centroid in the
    // material coordinates, RMS radius s and 1/s; the Galerkin blocks on the basis
    // phi = (1, (X - c) / s) of each component (9 x 9, upper triangle packed, 45
    // values) are summed per chunk of at most kGal9Rows hord positions of one
    // aggregate, then per aggregate, and factored (ml_setup).
    bool lin_built = false;       // the data below built for this hierarchy (on first use)
    std::vector<int> hord_h;      // host copy of hord
    bool lin_ok = false;          // levels >= 3 exist, hold at most kCoarse9WG aggregates, contiguous in hord
    bool lin_on = false;          // the last ml_setup formed the 9 x 9 blocks
    double *geo = nullptr;        // 4 per aggregate (all levels): c_u, c_v, s, 1/s
    int *gchunk = nullptr;        // 4 per chunk: level, global aggregate, first and end hord position
    int *gcp = nullptr;           // n_3 + ... + n_L + 1: chunk range of each aggregate of levels >= 3
    int nchunks = 0;
    double *gpart = nullptr;      // 45 per chunk
    double *D9 = nullptr;         // 54 per aggregate of levels >= 3: factors of the Galerkin block
    double *feat2 = nullptr;      // 2 per hord position: (phi_1, phi_2) of the vertex on its level-2 aggregate
    double *feat3 = nullptr;      // 2 per vertex: (phi_1, phi_2) on its level-3 aggregate
    double *R9p = nullptr;        // 9 per restriction workgroup (4 per level-3 aggregate): partial level-3 sums
    double *C9 = nullptr;         // 9 per aggregate of levels >= 3: accumulated coarse field
    std::vector<int> counts;      // host copy of n[0..L]
    std::vector<std::vector<int>> maps;   // host: level-l aggregate of every vertex (l = 1..L)
};

// Learned block preconditioner (message-passing network over the vertex graph).
constexpr int kLpFeat = 16, kLpHidden = 64, kLpOut = 6, kLpLayers = 3;
// flat parameter vector: W_in[16][64], b_in[64], {W_l[128][64], b_l[64]} x 3, W_out[64][6], b_out[6]
constexpr int kLpOffIn = 0;
constexpr int kLpLayerSize = 2 * kLpHidden * kLpHidden + kLpHidden;   // 8256
constexpr int kLpOffLayer0 = kLpFeat * kLpHidden + kLpHidden;          // 1088
constexpr int kLpOffOut = kLpOffLayer0 + kLpLayers * kLpLayerSize;     // 25856
constexpr int kLpParams = kLpOffOut + kLpHidden * kLpOut + kLpOut;     // 26246
// bfloat16 copies used by the matrix units (unsigned short): W_in^T[64][16],
// W_l^T[64][128] x 3 (forward), W_l[128][64] x 3 (backward), W_out[64][6]
constexpr int kLqInT = 0;
constexpr int kLqLayerT = kLpHidden * kLpFeat;                             // 1024
constexpr int kLqLayer = kLqLayerT + kLpLayers * 2 * kLpHidden * kLpHidden; // 25600
constexpr int kLqOut = kLqLayer + kLpLayers * 2 * kLpHidden * kLpHidden;    // 50176
constexpr int kLqSize = kLqOut + kLpHidden * kLpOut;                        // 50560
constexpr int kLpGradWG = 256;   // maximum workgroups of the weight-gradient kernels

struct LearnedState {
    bool have_weights = false;
    bool zero_output = false;   // output layer all zero: the blocks are the block-Jacobi ones
    bool with_ml = false;        // the training loss adds the multilevel coarse corrections to z
    int N = 0;                   // vertex count the buffers were sized for
    int64_t nbs = 0;
    int adam_t = 0;
    float *theta = nullptr, *grad = nullptr, *adam_m = nullptr, *adam_v = nullptr;   // kLpParams
    unsigned short *wq = nullptr;                                                     // kLqSize
    int *tpos = nullptr;         // nbs: SELL entry of the transposed block, -1 for padding
    float *w = nullptr, *wT = nullptr;   // nbs edge weights (row-normalised) and their transpose
    unsigned short *X = nullptr;         // N x 16
    unsigned short *h[kLpLayers + 1] = {nullptr, nullptr, nullptr, nullptr};   // N x 64
    unsigned short *a[kLpLayers] = {nullptr, nullptr, nullptr};                // N x 64
    unsigned short *dpre[2] = {nullptr, nullptr};                              // N x 64
    float *o = nullptr, *dO = nullptr;   // N x 6
    float *dinp = nullptr;               // N x 128
    float *gpart = nullptr;              // weight-gradient partials
    size_t gpart_n = 0;
    double *dpart = nullptr;             // FP64 per-workgroup partial sums (ceil(N / 256))
    double *scal = nullptr;              // 8: sum log t, r.e, z.r, z.Az, loss
    double *e = nullptr, *r = nullptr, *z = nullptr, *Az = nullptr;   // 3N training vectors
};

}  // namespace eolc

struct eolc_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int flags = 0;
    // material
    double rho = 0.1, E = 500.0, nu = 0.3, beta = 1e-5, dampA = 0.0, dampB = 1.0;
    bool have_topo = false, have_mat = false, have_state = false, have_forces = false;
    double h_last = 0.0;
    eolc::Topology topo;

    // ---- device buffers
    double *X = nullptr; int *tri = nullptr; int *hinge = nullptr;
    int *slice_off = nullptr, *srow = nullptr, *slen = nullptr, *scol = nullptr, *diagpos = nullptr;
    int *cptr = nullptr, *clist = nullptr, *fptr = nullptr, *flist = nullptr;
    double *stageK = nullptr;   // (6F + 10H) x 9
    double *stageF = nullptr;   // (3F + 4H) x 3
    double *stageM = nullptr;   // 6F mass scalars
    double *x = nullptr, *v = nullptr, *f = nullptr, *rhs = nullptr;
    double *A64 = nullptr;      // 9 x nbs (SoA, SELL entry order)
    float *S32 = nullptr;       // 9 x nbs: dampingB h^2 K_ij + m_ij I off the diagonal, 0 on it
    float *M32 = nullptr;       // N lumped row masses sum_j m_ij (FP32, written by k_rhs)
    // symmetric FP32 layout: upper records {column bits, S'_ij (9)} and lower records
    int *supper = nullptr, *uoff = nullptr, *loff = nullptr, *lrec = nullptr, *e2u = nullptr;
    int *wwin = nullptr;        // 2 per PCG workgroup: its vertex window (Topology::wwin)
    float *S32u = nullptr;      // 10 x nup
    float *S32a = nullptr;      // 10 x nbs: {column bits, S'_ij} per SELL entry (record layout)
    int s32_valid = 0;          // layouts (kLay* bits) holding the last assembly's values
    double *S64u = nullptr;     // 10 x nup: the symmetric records in FP64 (FP64 SpMV layout 4)
    // patch layout (Topology::pt_*) and its records {column bits, S'_ij (9)}
    int *pt_row = nullptr, *pt_cnt = nullptr, *pt_roff = nullptr, *pt_soff = nullptr, *pt_loff = nullptr;
    unsigned short *pt_lo = nullptr;
    int *pt_e2r = nullptr, *pt_e2l = nullptr, *pt_woff = nullptr, *pt_win = nullptr;
    unsigned char *stiff = nullptr;   // nbs: the SELL entry's block is applied in FP64
    double *A64s = nullptr;           // 9 x nbs: such a block's FP64 values, contiguous per entry
    unsigned long long *smask = nullptr;   // N: bit k = the row's SELL entry k is stiff
    int *sellbase = nullptr;          // N: SELL entry 0 of the row (entry k at + 64 k)
    float *P32 = nullptr;       // 10 x pt_nrec
    double *P64 = nullptr;      // 10 x pt_nrec
    double *M64 = nullptr;      // N row masses sum_j m_ij in FP64 (written by k_rhs)
    int *entry_is_diag = nullptr;  // nbs
    double *Mval = nullptr;     // nbs mass scalars (M block = m I3)
    double *D64 = nullptr;      // 6 x N  block-Jacobi inverse (symmetric, SoA)
    float *D32 = nullptr;
    unsigned char *fixmask = nullptr;  // N, bit k = DOF k fixed
    double *vfix = nullptr;     // 3N fixed values (0 where free)
    int *fix_dofs = nullptr;    // 3N: the current call's fixed DOF list (no per-call allocation)
    double *fix_vals = nullptr; // 3N: its values
    // PCG work vectors
    double *b64 = nullptr, *xs64 = nullptr, *r64 = nullptr, *z64 = nullptr, *q64 = nullptr;
    double *p64[2] = {nullptr, nullptr};
    float *r32 = nullptr, *z32 = nullptr, *q32 = nullptr, *xl32 = nullptr;   // xl32: the batched solve's
    double *xl64 = nullptr;     // the mixed solve's FP64 correction since the last reliable update
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
    std::vector<int> topo_hinge;
    eolc::EolState eol;
    double *cd_unit = nullptr;  // 3N jitter values in (-1, 1)
    int cd_host_decisions = 0;  // angle tests re-decided on the host (diagnostic)
    // learned preconditioner
    eolc::LearnedState lp;
    int spmv_variant[2] = {0, 0};   // SpMV launch variant for FP32 [0] and FP64 [1] solves
    // FP32 patch records: a block whose largest entry exceeds stiff_rel x dampingB h^2 E
    // is applied from its FP64 values by both its rows (Topology::sellbase, stiff masks)
    double stiff_rel = 10.0;
    double e_batch_max = 0.0;       // largest per-instance Young's modulus of a batch
    double mixed_delta = 0.5;       // reliable-update trigger of the mixed solves
    double mixed_delta_ml = 0.2;    // the same with the multilevel levels (shorter cycles)
    int ml_linear = 1;              // linear coarse fields: 0 never, 1 in FP32 iterations, 2 always
    double batch_delta = 0.1;       // the batched solve (its FP64 residuals cost more per iteration)
    // per-element material (batched cloths): Young's modulus per face, bending
    // stiffness per hinge; null = the context's scalar material
    double *e_face = nullptr;
    double *beta_hinge = nullptr;
    eolc::BatchState batch;
    int bj_iters_hint = 0;      // iterations of the last converged block-Jacobi solve
    int bj_iters_hint_ml = 0;   // the same with the multilevel preconditioner
    int fp64_sticky = 0;        // mixed solves left that start in FP64 after an FP32 stall
    int stall_first = 10000;    // FP32 stall limits (see kStallMinCycle) and the sticky count
    int stall_min = 1000;
    int stall_sticky = 8;
    eolc::MLState ml;           // aggregation hierarchy (built on first use per topology)
    bool ml_on = false;         // the current solve applies the multilevel levels
};

namespace eolc {
// Every ABI function body is a function-try-block whose handler calls this: no
// exception crosses the C ABI.  A failed host allocation reports
// EOLC_ERR_NO_MEMORY, anything else EOLC_ERR_INVALID; the message goes to c->err
// when there is a context.
inline int report_exception(eolc_ctx *c, const char *who) noexcept
{
    int code = EOLC_ERR_INVALID;
    const char *what = "unexpected exception";
    try {
        throw;
    } catch (const std::bad_alloc &) {
        code = EOLC_ERR_NO_MEMORY;
        what = "out of host memory";
    } catch (const std::exception &e) {
        what = e.what();
    } catch (...) {
    }
    if (c) {
        try {
            c->err = std::string(who) + ": " + what;
        } catch (...) {
        }
    }
    return code;
}

// kernels (k_elements.hip, k_assemble.hip, k_pcg.hip)
hipError_t launch_face_eval(eolc_ctx *c, double h, const double *grav);
hipError_t launch_hinge_eval(eolc_ctx *c, double h);
// FP32 matrix layouts (bits of launch_assemble's `lay`): SoA S32 (SpMV layouts 0 and 3,
// the batched solve), symmetric upper records S32u (layout 1), per-entry records
// S32a (layout 2)
constexpr int kLaySoA = 1, kLaySym = 2, kLayRecord = 4;
constexpr int kLaySym64 = 8;                            // symmetric records in FP64 (S64u, FP64 layout 4)
constexpr int kLayPatch = 16, kLayPatch64 = 32;          // patch records in FP32 (P32) and FP64 (P64), layout 5
int fp32_layout(int variant);                          // the layout bit a FP32 SpMV variant reads
int fp64_layout(int variant);                          // the extra layout bit a FP64 variant reads (or 0)
hipError_t launch_assemble(eolc_ctx *c, int lay);      // FP64 values + the layouts in lay
hipError_t ensure_fp32_layout(eolc_ctx *c, int variant);
hipError_t ensure_fp64_layout(eolc_ctx *c, int variant);
hipError_t launch_rhs(eolc_ctx *c, double h);
hipError_t launch_update_x(eolc_ctx *c, double h);

int pcg_grid(int N);
int pcg_grid_phys(int G);   // physical workgroups of the row kernels with a second logical one
hipError_t pcg_setup(eolc_ctx *c, bool mixed, int precond_kind, bool ml_linear = true);
hipError_t pcg_init(eolc_ctx *c, bool mixed);
hipError_t pcg_chunk(eolc_ctx *c, bool mixed, int k0, int iters);
hipError_t pcg_reliable_update(eolc_ctx *c);
hipError_t pcg_restart_from_guess(eolc_ctx *c);   // a stalled mixed solve: FP64 from the initial guess
hipError_t pcg_restart_inner(eolc_ctx *c, bool mixed, int k);
hipError_t pcg_precondition(eolc_ctx *c, bool mixed);   // z = P r for r in r64 (parity 0)
hipError_t pcg_replace_residual64(eolc_ctx *c);         // FP64 mode: r = b - A x, ||r||^2 -> rmax2
hipError_t pcg_finish(eolc_ctx *c, double *out);        // solution with fixed values -> out
hipError_t spmv_true_residual(eolc_ctx *c, const double *v);   // ||rhs - A v||^2 -> true_rr
hipError_t scal_set(eolc_ctx *c, double tol2, double delta2, int reset);
hipError_t scal_delta(eolc_ctx *c, double delta2);      // reliable-update trigger only
hipError_t set_fixed(eolc_ctx *c, int n, const int *d_dofs, const double *d_vals, int set);
// d_vals[t] = (1 - 0.01) v[d_dofs[t]] + d_vals[t] on the context's stream.
hipError_t fixed_from_vold(eolc_ctx *c, int n, const int *d_dofs, double *d_vals);
// EOL rows of the last forces_fill (k_eol.hip); after launch_rhs.
hipError_t eol_fill(eolc_ctx *c);
constexpr int kSpmvVariants = 29;
// LDS-staged SpMV variants: a workgroup's vertex window holds at most this many
// vertices (3 x 5440 values of p: 65,280 B in FP32, 130,560 B in FP64)
constexpr int kStageVerts = 5440;
bool spmv_staged(int variant);   // the variant gathers p from LDS
int spmv_unstaged(int variant);  // a variant without staging that reads the same layout
bool spmv_patch(int variant);    // the variant reads the patch layout
// the variant can run on this topology (staged: windows within kStageVerts; patch:
// the patch layout was built), else spmv_fallback(variant) gives the same bits
bool spmv_usable(const Topology &t, int variant);
int spmv_fallback(int variant);
hipError_t time_spmv(eolc_ctx *c, bool mixed, int variant, int reps, float *ms);
hipError_t stream_copy(eolc_ctx *c, size_t bytes, int reps, float *ms, int read_only);
// learned preconditioner (k_learned.hip); buffers allocated by the API layer
void learned_wgrad_shape(int N, int *G, int *rows);   // weight-gradient workgroups, rows per workgroup
hipError_t learned_quantize(eolc_ctx *c);
hipError_t learned_forward(eolc_ctx *c);
hipError_t learned_precond(eolc_ctx *c);               // forward + P -> D32 (fixmask set)
hipError_t learned_loss_grad(eolc_ctx *c);             // lp.e set and masked; loss -> lp.scal[4]
hipError_t learned_adam(eolc_ctx *c, float lr);
hipError_t learned_mask_error(eolc_ctx *c);            // lp.e <- e with fixed DOFs zeroed
hipError_t scal_restart(eolc_ctx *c, int k);           // iteration k starts with beta = 0
// multilevel preconditioner (ml_host.cpp builds the hierarchy, k_ml.hip applies it)
std::string ml_build(const Topology &t, std::vector<std::vector<int>> &agg_levels, std::vector<int> &hord,
                     std::vector<int> &ptr1, std::vector<int> &agg1, std::vector<int> &cptr, std::vector<int> &par,
                     std::vector<unsigned char> &lam, MLLevels &lv);
hipError_t ml_setup(eolc_ctx *c, bool linear = true);         // Galerkin diagonal blocks of every level
hipError_t ml_apply(eolc_ctx *c, bool mixed, int parity, int check_done);   // z = D0 r + coarse, r.z partials
hipError_t ml_coarse64(eolc_ctx *c);                          // z64 = coarse corrections of r64 (no fine term)
// velocity transfer after a remesh (k_transfer.hip); host arrays in and out
hipError_t velocity_transfer(eolc_ctx *c, int n_old, const double *u_old, const double *v_old, int f_old,
                             const int *tri_old, int n_new, const double *u_new, double *v_new, int *src, int *face,
                             double *bary, int *n_missing, float *ms, const double *x_old = nullptr,
                             const int *eol_old = nullptr, int n_eol = 0, const double *V_old = nullptr);
// batched cloths (k_batch.hip)
hipError_t batch_set_fixmask(eolc_ctx *c, const int *d_nodes, int n);
hipError_t batch_solve_async(eolc_ctx *c, double tol, int maxit, int variant);
void build_collision_edges(eolc_ctx *c, int F, const int *tri);
// scratch of eolc_collide_noalloc (k_collide.hip): the device address of this GPU's
// array (hipGetSymbolAddress once per device, with that device current), and the
// lock that serialises the calls that use it within the process
constexpr size_t kCollideScratchBytes = (size_t)256 << 20;
hipError_t collide_scratch(int device, void **ptr);
std::mutex *collide_scratch_lock(int device);
void build_jitter(std::vector<double> &u, int N);
}  // namespace eolc
