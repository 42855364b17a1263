/*
 * orc_mesh.c -- synthetic cloth grid and hinge (interior edge) list (oracle).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * orc_grid_mesh: n x n quads over the material square [0,1]^2, vertex
 *   (i, j) -> index j*(n+1)+i, X = (i/n, j/n).  Quad (i, j) is split along
 *   v00-v11 when (i+j) is even and along v10-v01 otherwise (checkerboard), both
 *   triangles counter-clockwise in material space (positive signed area, which
 *   ComputeMembrane.cpp:43-44 and ComputeInertial.cpp:33 rely on).
 *
 * orc_build_hinges: restates how ArcSim creates and orients edges
 *   (external/ArcSim/mesh.cpp:356-375): faces are added in order; for each face
 *   the edges (v[i], v[i+1]) are created in order i = 0,1,2 when missing, with
 *   n[0] = v[i], n[1] = v[i+1]; face slot side = 0 if the face traverses
 *   n[0] -> n[1], else 1.  An edge is a hinge when both sides are set
 *   (Forces.cpp:688-690).  Hinge = (n[0], n[1], apex of adjf[0], apex of adjf[1])
 *   (Forces.cpp:692-697), in edge-creation order.
 *
 * Parity: pinned by tests of the counts (interior edge count = 3F/2 - boundary/2),
 * orientation (each hinge's apexes lie on opposite sides, adjf[0] traverses
 * x0 -> x1) and a brute-force pair search on small grids.
 */
#include "oracle.h"
#include <stdlib.h>

int orc_grid_mesh(int n, double *X, int *tri)
{
    int np = n + 1;
    for (int j = 0; j <= n; ++j)
        for (int i = 0; i <= n; ++i) {
            X[2 * (j * np + i)] = (double)i / n;
            X[2 * (j * np + i) + 1] = (double)j / n;
        }
    int f = 0;
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            int v00 = j * np + i, v10 = v00 + 1, v01 = v00 + np, v11 = v01 + 1;
            if (((i + j) & 1) == 0) {
                tri[3 * f] = v00; tri[3 * f + 1] = v10; tri[3 * f + 2] = v11; ++f;
                tri[3 * f] = v00; tri[3 * f + 1] = v11; tri[3 * f + 2] = v01; ++f;
            } else {
                tri[3 * f] = v00; tri[3 * f + 1] = v10; tri[3 * f + 2] = v01; ++f;
                tri[3 * f] = v10; tri[3 * f + 1] = v11; tri[3 * f + 2] = v01; ++f;
            }
        }
    return f;
}

typedef struct { int64_t key; int id; } keyid;
static int cmp_keyid(const void *a, const void *b)
{
    const keyid *x = (const keyid *)a, *y = (const keyid *)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->id < y->id ? -1 : (x->id > y->id);
}

int orc_build_hinges(int N, int F, const int *tri, int *hinge, int cap)
{
    /* one record per (face, local edge) in creation order; group by vertex pair */
    int64_t m = 3 * (int64_t)F;
    keyid *ks = (keyid *)malloc(sizeof(keyid) * m);
    for (int64_t r = 0; r < m; ++r) {
        int f = (int)(r / 3), i = (int)(r % 3);
        int a = tri[3 * f + i], b = tri[3 * f + (i + 1) % 3];
        int lo = a < b ? a : b, hi = a < b ? b : a;
        ks[r].key = (int64_t)lo * N + hi;
        ks[r].id = (int)r;
    }
    qsort(ks, m, sizeof(keyid), cmp_keyid);
    /* edge record: created by the first face-edge (lowest r) in each group */
    int *first = (int *)malloc(sizeof(int) * m);   /* group leader per r */
    for (int64_t k = 0; k < m;) {
        int64_t e = k;
        while (e < m && ks[e].key == ks[k].key) ++e;
        for (int64_t q = k; q < e; ++q) first[ks[q].id] = ks[k].id;
        k = e;
    }
    /* adjf[side] per edge leader; later faces overwrite like ArcSim's e->adjf[side] = face */
    int *adj0 = (int *)malloc(sizeof(int) * m), *adj1 = (int *)malloc(sizeof(int) * m);
    for (int64_t r = 0; r < m; ++r) adj0[r] = adj1[r] = -1;
    for (int64_t r = 0; r < m; ++r) {
        int f = (int)(r / 3), i = (int)(r % 3);
        int a = tri[3 * f + i];
        int L = first[r];
        int n0 = tri[3 * (L / 3) + L % 3];
        if (a == n0) adj0[L] = f; else adj1[L] = f;
    }
    int H = 0;
    for (int64_t r = 0; r < m; ++r) {
        if (first[r] != r) continue;                    /* not an edge creator */
        if (adj0[r] < 0 || adj1[r] < 0) continue;       /* boundary */
        int f = (int)(r / 3), i = (int)(r % 3);
        int n0 = tri[3 * f + i], n1 = tri[3 * f + (i + 1) % 3];
        int fa = adj0[r], fb = adj1[r];
        int x2 = -1, x3 = -1;
        for (int k = 0; k < 3; ++k) {
            int p = tri[3 * fa + k], q = tri[3 * fb + k];
            if (p != n0 && p != n1) x2 = p;
            if (q != n0 && q != n1) x3 = q;
        }
        if (H < cap) {
            hinge[4 * H] = n0; hinge[4 * H + 1] = n1; hinge[4 * H + 2] = x2; hinge[4 * H + 3] = x3;
        }
        ++H;
    }
    free(ks); free(first); free(adj0); free(adj1);
    return H;
}
