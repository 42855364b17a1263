// Host C++ step driver on the drop-in classes: the Cloth::step loop of the
// reference (Cloth.cpp:359-413) for a square n x n cloth with its two y = 1
// corners fixed, remeshing and EOL off, solver "none".
//
//   cloth_step <n> <steps> <out.bin>
//
// Writes the final positions (3N doubles) to out.bin and prints one summary line.
#include <cstdio>
#include <cstdlib>
#include <map>
#include <utility>

#include "eolc_dropin.hpp"

namespace {

// n x n quads over [0,1]^2, checkerboard diagonals, counter-clockwise faces
void grid(int n, eolc::MeshArrays &m)
{
    const int np = n + 1, N = np * np;
    m.X.resize(2 * (size_t)N);
    m.x.assign(3 * (size_t)N, 0.0);
    m.v.assign(3 * (size_t)N, 0.0);
    for (int j = 0; j <= n; ++j)
        for (int i = 0; i <= n; ++i) {
            const int k = j * np + i;
            m.X[2 * k] = m.x[3 * k] = (double)i / n;
            m.X[2 * k + 1] = m.x[3 * k + 1] = (double)j / n;
        }
    m.tri.clear();
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            const int v00 = j * np + i, v10 = v00 + 1, v01 = v00 + np, v11 = v01 + 1;
            if (((i + j) & 1) == 0) m.tri.insert(m.tri.end(), {v00, v10, v11, v00, v11, v01});
            else m.tri.insert(m.tri.end(), {v00, v10, v01, v10, v11, v01});
        }
}

// interior edges in ArcSim's creation order and orientation (mesh.cpp:356-375)
void hinges(eolc::MeshArrays &m)
{
    struct EdgeRec { int n0, n1, adj[2], order; };
    std::map<std::pair<int, int>, EdgeRec> edges;
    const int F = (int)(m.tri.size() / 3);
    int order = 0;
    for (int f = 0; f < F; ++f)
        for (int i = 0; i < 3; ++i) {
            const int a = m.tri[3 * f + i], b = m.tri[3 * f + (i + 1) % 3];
            auto key = std::make_pair(std::min(a, b), std::max(a, b));
            auto it = edges.find(key);
            if (it == edges.end()) it = edges.emplace(key, EdgeRec{a, b, {-1, -1}, order++}).first;
            it->second.adj[a == it->second.n0 ? 0 : 1] = f;
        }
    std::vector<const EdgeRec *> byorder(edges.size());
    for (auto &kv : edges) byorder[kv.second.order] = &kv.second;
    m.hinge.clear();
    for (const EdgeRec *e : byorder) {
        if (e->adj[0] < 0 || e->adj[1] < 0) continue;
        int apex[2];
        for (int s = 0; s < 2; ++s)
            for (int k = 0; k < 3; ++k) {
                const int v = m.tri[3 * e->adj[s] + k];
                if (v != e->n0 && v != e->n1) apex[s] = v;
            }
        m.hinge.insert(m.hinge.end(), {e->n0, e->n1, apex[0], apex[1]});
    }
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc < 4) {
        std::fprintf(stderr, "usage: cloth_step <n> <steps> <out.bin>\n");
        return 2;
    }
    const int n = std::atoi(argv[1]), steps = std::atoi(argv[2]);
    try {
        eolc::MeshArrays mesh;
        grid(n, mesh);
        hinges(mesh);
        eolc::Material mat;
        mat.density = 0.1; mat.e = 500.0; mat.nu = 0.3; mat.beta = 1e-5;
        const double grav[3] = {0.0, 0.0, -9.8}, h = 5e-3;
        const int fixedNodes[2] = {n * (n + 1), (n + 1) * (n + 1) - 1};
        eolc::Device dev(0);
        dev.setTopology(mesh);
        eolc::Forces forces(dev);
        eolc::GeneralizedSolver solver(dev);
        solver.tol = 1e-12;
        int iters = 0;
        for (int s = 0; s < steps; ++s) {
            forces.fill(mesh, mat, grav, h, false);
            std::vector<int> dofs;
            std::vector<double> vals;
            for (int node : fixedNodes)
                for (int k = 0; k < 3; ++k) {
                    dofs.push_back(3 * node + k);
                    vals.push_back((1 - 0.01) * mesh.v[3 * node + k]);   // Constraints.cpp:117
                }
            if (!solver.velocitySolve(dofs, vals, mesh.v)) {
                std::fprintf(stderr, "solve failed: %s\n", eolc_last_error(dev.ctx()));
                return 1;
            }
            iters += solver.stats.iters;
            for (size_t i = 0; i < mesh.x.size(); ++i) mesh.x[i] = mesh.x[i] + h * mesh.v[i];   // Cloth.cpp:394-410
        }
        FILE *fp = std::fopen(argv[3], "wb");
        if (!fp) return 1;
        std::fwrite(mesh.x.data(), sizeof(double), mesh.x.size(), fp);
        std::fclose(fp);
        std::printf("cloth_step n=%d steps=%d hinges=%zu iterations=%d\n", n, steps, mesh.hinge.size() / 4, iters);
    } catch (const eolc::Error &e) {
        std::fprintf(stderr, "eolc error %d: %s\n", e.code, e.what());
        return 1;
    }
    return 0;
}
