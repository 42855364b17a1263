// Constraints::fill, collision part (Constraints.cpp:417-470): inequality rows of the
// velocity QP from the proximity records of CD2.  Host code; the records come from
// the GPU collision kernels (eolc_collide) and the rows go to the host QP.
//
// Row per record, in record order:
//   (count1, count2) = (3, 1)  cloth vertex vs box face:  -nor1 on verts2[0]
//   (2, 2)                     edge vs edge:              -nor2 * weights2[j] on verts2[j], j < 2
//   (1, 3)                     point / box corner vs face: -nor2 * weights2[j] on verts2[j], j < 3
// other records give no row, and a row touching an EOL node is skipped.  The
// right-hand side bineq is zero.  The matrix is returned as scalar CSR with the
// columns of each row increasing and duplicate entries summed -- the content of the
// Eigen matrix setFromTriplets builds.
#include <algorithm>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "eolc_internal.h"

extern "C" int eolc_collision_constraints(eolc_ctx *c, const eolc_collision *recs, int n, int N,
                                          const unsigned char *eol, int64_t *rowptr, int *col, double *val,
                                          int cap_rows, int64_t cap_nnz, int *rows, int64_t *nnz)
try {
    auto fail = [&](int code, const char *msg) {
        if (c) c->err = msg;
        return code;
    };
    if (n < 0 || N <= 0 || (n && !recs) || !rows || !nnz) return fail(EOLC_ERR_INVALID, "collision_constraints: bad arguments");
    std::vector<std::vector<std::pair<int, double>>> R;
    for (int r = 0; r < n; ++r) {
        const eolc_collision &cl = recs[r];
        int nv = 0;
        const double *nor = nullptr;
        bool use_w = true;
        if (cl.count1 == 3 && cl.count2 == 1) {
            nv = 1;
            nor = cl.nor1;
            use_w = false;
        } else if (cl.count1 == 2 && cl.count2 == 2) {
            nv = 2;
            nor = cl.nor2;
        } else if (cl.count1 == 1 && cl.count2 == 3) {
            nv = 3;
            nor = cl.nor2;
        } else {
            continue;
        }
        bool skip = false;
        for (int j = 0; j < nv; ++j) {
            const int v = cl.verts2[j];
            if (v < 0 || v >= N) return fail(EOLC_ERR_INVALID, "collision_constraints: record vertex out of range");
            if (eol && eol[v]) skip = true;
        }
        if (skip) continue;
        std::vector<std::pair<int, double>> row;
        for (int j = 0; j < nv; ++j)
            for (int k = 0; k < 3; ++k)
                row.emplace_back(3 * cl.verts2[j] + k, use_w ? -nor[k] * cl.weights2[j] : -nor[k]);
        std::stable_sort(row.begin(), row.end(),
                         [](const std::pair<int, double> &a, const std::pair<int, double> &b) { return a.first < b.first; });
        std::vector<std::pair<int, double>> merged;
        for (const auto &e : row) {
            if (!merged.empty() && merged.back().first == e.first) merged.back().second += e.second;
            else merged.push_back(e);
        }
        R.push_back(std::move(merged));
    }
    int64_t total = 0;
    for (const auto &row : R) total += (int64_t)row.size();
    *rows = (int)R.size();
    *nnz = total;
    if ((int)R.size() > cap_rows || total > cap_nnz || (!R.empty() && (!rowptr || !col || !val)))
        return fail(EOLC_ERR_OVERFLOW, "collision_constraints: output capacity too small");
    if (rowptr) {
        int64_t k = 0;
        rowptr[0] = 0;
        for (size_t r = 0; r < R.size(); ++r) {
            for (const auto &e : R[r]) {
                col[k] = e.first;
                val[k] = e.second;
                ++k;
            }
            rowptr[r + 1] = k;
        }
    }
    return EOLC_OK;
} catch (...) {
    return eolc::report_exception(c, "collision_constraints");
}
