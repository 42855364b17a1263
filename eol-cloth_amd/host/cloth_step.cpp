// Host C++ step driver on the drop-in classes: the Cloth::step loop of the
// reference (Cloth.cpp:359-413) for a square n x n cloth with its two y = 1
// corners fixed, remeshing and EOL off, solver "none".
//
//   cloth_step <n> <steps> <out.bin> [scene]
//
// Writes the final positions (3N doubles) to out.bin and prints one summary line.
// With "scene", the steps run through eolc::Scene (Scene.cpp:72-98) instead: the
// two corners start fixed and, from t > 0.0475 s, move with scripted velocities
// (one corner x = 0.1 m/s on all axes, the other only its z axis scripted at
// 0.05 m/s, as two FixedLists); a box rises and turns under the cloth and two
// static points sit near it, with CD2 on every step.  out.bin then also holds the
// box frame (16 doubles, column-major) after the positions, and out.bin.cd the
// CD2 records (eolc_collision) of the final state.
#include <cstdio>
#include <cstdlib>
#include <map>
#include <string>
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

int run_scene(int n, int steps, const char *out)
{
    try {
        eolc::MeshArrays mesh;
        grid(n, mesh);
        hinges(mesh);
        eolc::Device dev(0);
        dev.setTopology(mesh);
        dev.check(eolc_set_material(dev.ctx(), 0.1, 500.0, 0.3, 1e-5, 0.0, 1.0));
        dev.check(eolc_upload_state(dev.ctx(), mesh.x.data(), mesh.v.data()));
        eolc::Scene s(dev);
        s.tol = 1e-12;
        s.threshold = 5e-3;
        const int cA = n * (n + 1), cB = (n + 1) * (n + 1) - 1;   // the y = 1 corners
        eolc::FixedList f0, f1;
        f0.ci[0] = f1.ci[0] = cA;
        f0.ci[1] = f1.ci[1] = cB;
        for (int a = 0; a < 3; ++a) f0.c[0][a] = f0.c[1][a] = f1.c[0][a] = 1.0;
        for (int a = 3; a < 6; ++a) f0.c[0][a] = f0.c[1][a] = f1.c[0][a] = 0.0;
        f1.when = 0.0475;
        f1.c[0][3] = 0.1;                                         // corner A: x scripted
        const double cBspec[6] = {0.0, 0.0, 1.0, 0.0, 0.0, 0.05};  // corner B: z only
        for (int a = 0; a < 6; ++a) f1.c[1][a] = cBspec[a];
        s.fs = {f0, f1};
        eolc::Box b;
        b.dim[0] = 0.4; b.dim[1] = 0.3; b.dim[2] = 0.1;
        b.E[12] = 0.5; b.E[13] = 0.45; b.E[14] = -0.08;           // centre
        const double bv[6] = {0.0, 0.0, 2.0, 0.0, 0.0, 0.4};      // turns about z, rises
        for (int k = 0; k < 6; ++k) b.v[k] = bv[k];
        s.boxes = {b};
        s.pts = {0.3, 0.3, -0.004, 0.7, 0.6, 0.003};
        s.nrm = {0.0, 0.0, 1.0, 0.0, 0.0, -1.0};
        int iters = 0, switched = -1;
        size_t records = 0;
        for (int k = 0; k < steps; ++k) {
            const size_t before = s.fsindex;
            s.step();
            if (s.fsindex != before) switched = k;
            iters += s.stats.iters;
            records += s.cls.size();
        }
        s.detect();   // the CD2 of the final state (what the next step would hand the QP)
        std::vector<double> x(mesh.x.size()), v(mesh.x.size());
        dev.check(eolc_download_state(dev.ctx(), x.data(), v.data()));
        FILE *fp = std::fopen(out, "wb");
        if (!fp) return 1;
        std::fwrite(x.data(), sizeof(double), x.size(), fp);
        std::fwrite(s.boxes[0].E, sizeof(double), 16, fp);
        std::fclose(fp);
        fp = std::fopen((std::string(out) + ".cd").c_str(), "wb");
        if (!fp) return 1;
        std::fwrite(s.cls.data(), sizeof(eolc_collision), s.cls.size(), fp);
        std::fclose(fp);
        std::printf("cloth_step scene n=%d steps=%d switch_step=%d iterations=%d records_per_run=%zu final_records=%zu\n",
                    n, steps, switched, iters, records, s.cls.size());
    } catch (const eolc::Error &e) {
        std::fprintf(stderr, "eolc error %d: %s\n", e.code, e.what());
        return 1;
    }
    return 0;
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc < 4) {
        std::fprintf(stderr, "usage: cloth_step <n> <steps> <out.bin> [scene]\n");
        return 2;
    }
    const int n = std::atoi(argv[1]), steps = std::atoi(argv[2]);
    if (argc > 4 && std::string(argv[4]) == "scene") return run_scene(n, steps, argv[3]);
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
