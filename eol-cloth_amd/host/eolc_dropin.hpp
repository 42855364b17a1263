// Drop-in C++ classes with the reference's interfaces for solver "none", built
// on the eolc C ABI.  Header-only; link against libeolc.so.
//
//   eolc::Forces             Forces.h:26-45   -- fill(mesh, material, grav, h) -> f, M, MDK
//   eolc::GeneralizedSolver  GeneralizedSolver.h:12-32 -- velocitySolve(...) for solver "none"
//   eolc::CD / eolc::CD2     Collisions.h:9-11 -- proximity queries into btc::Collision-like records
//   eolc::fillConstraints    Constraints.cpp:114-119, 417-514 -- A_ineq rows from the records, fixed rows
//   eolc::FixedList, eolc::Box, eolc::Scene  FixedList.h, Box.cpp:103-114, Scene.cpp:72-98 --
//                            the scene step on the device-resident state (scripted corners, obstacle motion)
//
// The reference's types (ArcSim Mesh, Eigen vectors and sparse matrices) are not
// available here, so the mesh is passed as the flat arrays those types hold
// (MeshArrays) and matrices come back as scalar CSR (CsrMatrix).  A reference
// build replaces the bodies of its Forces::fill / velocitySolve / CD2 by calls to
// these classes after copying Node::x, Vert::u, Face::v[] and Edge::n[]/adjf[] into
// a MeshArrays (flat x, X, tri and hinge arrays in ArcSim order).
#pragma once
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/eolc.h"

namespace eolc {

struct Material {           // Cloth.h:28-35
    double density = 0.05, e = 50.0, nu = 0.01, beta = 1e-5, dampingA = 0.0, dampingB = 1.0;
};

struct MeshArrays {         // what Forces::fill reads from the ArcSim mesh
    std::vector<double> x;  // 3N  Node::x
    std::vector<double> v;  // 3N  Node::v
    std::vector<double> X;  // 2N  Vert::u (material space)
    std::vector<int> tri;   // 3F  Face::v[k]->node->index
    std::vector<int> hinge; // 4H  (n[0], n[1], apex of adjf[0], apex of adjf[1]) of interior edges
    int N() const { return (int)(x.size() / 3); }
};

struct CsrMatrix {
    int n = 0;
    std::vector<int64_t> rowptr;
    std::vector<int> col;
    std::vector<double> val;
};

class Error : public std::runtime_error {
public:
    Error(int code, const std::string &m) : std::runtime_error(m), code(code) {}
    int code;
};

// One GPU context shared by the classes below (one per host thread).
class Device {
public:
    explicit Device(int device = 0)
    {
        int rc = eolc_create(device, &ctx_);
        if (rc != EOLC_OK) throw Error(rc, "eolc_create failed");
    }
    ~Device() { eolc_destroy(ctx_); }
    Device(const Device &) = delete;
    Device &operator=(const Device &) = delete;
    eolc_ctx *ctx() const { return ctx_; }
    void check(int rc) const
    {
        if (rc != EOLC_OK) throw Error(rc, eolc_last_error(ctx_));
    }
    // call after every remesh (topology change); cheap to skip otherwise
    void setTopology(const MeshArrays &m)
    {
        check(eolc_set_topology(ctx_, m.N(), m.X.data(), (int)(m.tri.size() / 3), m.tri.data(),
                                (int)(m.hinge.size() / 4), m.hinge.data()));
        nEol_ = 0;
    }
    // EOL nodes in EoL_index order (Node::EoL, Node::EoL_index); after setTopology
    void setEol(const std::vector<int> &nodes)
    {
        check(eolc_set_eol(ctx_, (int)nodes.size(), nodes.data()));
        nEol_ = (int)nodes.size();
    }
    int eolCount() const { return nEol_; }

private:
    eolc_ctx *ctx_ = nullptr;
    int nEol_ = 0;
};

// Forces: f, M and MDK = M + dampingB h^2 K of the current state (Forces.cpp:912-930),
// with 3N + 2 EoL_Count rows when EOL nodes are set (Device::setEol).
class Forces {
public:
    explicit Forces(Device &d) : dev_(d) {}
    std::vector<double> f;
    CsrMatrix M, MDK;
    int EoL_cutoff = 0;

    void fill(const MeshArrays &mesh, const Material &mat, const double grav[3], double h, bool downloadMatrices = true)
    {
        dev_.check(eolc_set_material(dev_.ctx(), mat.density, mat.e, mat.nu, mat.beta, mat.dampingA, mat.dampingB));
        dev_.check(eolc_upload_state(dev_.ctx(), mesh.x.data(), mesh.v.data()));
        dev_.check(eolc_forces_fill(dev_.ctx(), grav, h));
        f.resize(mesh.x.size() + 2 * (size_t)dev_.eolCount());   // 3N + 2 EoL_Count
        dev_.check(eolc_download_forces(dev_.ctx(), f.data()));
        EoL_cutoff = (int)mesh.x.size();
        if (downloadMatrices) {
            download(EOLC_MAT_M, M);
            download(EOLC_MAT_MDK, MDK);
        }
    }

private:
    void download(int which, CsrMatrix &A)
    {
        int64_t nnz = 0;
        int rc = eolc_download_csr(dev_.ctx(), which, nullptr, nullptr, nullptr, 0, &nnz);
        if (rc != EOLC_ERR_OVERFLOW && rc != EOLC_OK) dev_.check(rc);
        A.n = (int)f.size();
        A.rowptr.resize(A.n + 1);
        A.col.resize(nnz);
        A.val.resize(nnz);
        dev_.check(eolc_download_csr(dev_.ctx(), which, A.rowptr.data(), A.col.data(), A.val.data(), nnz, &nnz));
    }
    Device &dev_;
};

// GeneralizedSolver::velocitySolve for solver "none" (GeneralizedSolver.cpp:107-172):
// MDK v = M v_old + h f with v[fixed] = 0.99 v_old[fixed] + scripted (the QP's
// single-DOF equality rows, Constraints.cpp:114-119).  Uses the MDK, M and f of the
// last Forces::fill on the same Device; v holds v_old on entry and the solution on exit.
class GeneralizedSolver {
public:
    explicit GeneralizedSolver(Device &d) : dev_(d) {}
    int mode = EOLC_SOLVE_FP64_BJ;
    double tol = 1e-8;
    int maxit = 100000;
    eolc_solve_stats stats{};

    bool velocitySolve(const std::vector<int> &fixedDofs, const std::vector<double> &fixedVals, std::vector<double> &v)
    {
        int rc = eolc_velocity_solve(dev_.ctx(), (int)fixedDofs.size(), fixedDofs.data(), fixedVals.data(), tol, maxit,
                                     mode, &stats);
        std::vector<double> x(v.size());
        dev_.check(eolc_download_state(dev_.ctx(), x.data(), v.data()));
        return rc == EOLC_OK;
    }

private:
    Device &dev_;
};

struct Obstacles {          // Obstacles / Points / Box (Points.h, Box.h)
    double threshold = 5e-3;
    std::vector<double> pts, nrm;   // 3P each
    std::vector<double> whd;        // 3B box dimensions
    std::vector<double> E;          // 16B column-major 4x4 frames
};

// CD2 (Collisions.cpp:55-77) and CD (Collisions.cpp:11-53): append to cls.
inline void collide(Device &dev, const Obstacles &obs, bool eolPoints, std::vector<eolc_collision> &cls)
{
    const int P = (int)(obs.pts.size() / 3), B = (int)(obs.whd.size() / 3);
    int count = 0;
    std::vector<eolc_collision> out(1024);
    for (;;) {
        int rc = eolc_collide(dev.ctx(), obs.threshold, P, obs.pts.data(), obs.nrm.data(), B, obs.whd.data(),
                              obs.E.data(), eolPoints ? 1 : 0, 8, 12, out.data(), (int)out.size(), &count);
        if (rc == EOLC_ERR_OVERFLOW && count > (int)out.size()) {
            out.resize(count);
            continue;
        }
        dev.check(rc);
        break;
    }
    cls.insert(cls.end(), out.begin(), out.begin() + count);
}
inline void CD2(Device &dev, const Obstacles &obs, std::vector<eolc_collision> &cls) { collide(dev, obs, false, cls); }
inline void CD(Device &dev, const Obstacles &obs, std::vector<eolc_collision> &cls) { collide(dev, obs, true, cls); }

// Constraints::fill (Constraints.cpp:122-514) without the EOL rows: the collision
// inequality rows from CD2 records (A_ineq v >= 0, b_ineq = 0) and the fixed-DOF
// equality rows of addFixed (Constraints.cpp:114-119): one row per fixed DOF with
// coefficient 1 and b = (1 - 0.01) v + scripted velocity.
struct QpConstraints {
    CsrMatrix Aineq, Aeq;
    std::vector<double> bineq, beq;
};

inline QpConstraints fillConstraints(const std::vector<eolc_collision> &cls, int N, const std::vector<int> &fixedNodes,
                                     const std::vector<double> &v, const std::vector<double> &scripted = {},
                                     const std::vector<unsigned char> &eol = {})
{
    QpConstraints q;
    int rows = 0;
    int64_t nnz = 0;
    const unsigned char *e = eol.empty() ? nullptr : eol.data();
    int rc = eolc_collision_constraints(nullptr, cls.data(), (int)cls.size(), N, e, nullptr, nullptr, nullptr, 0, 0,
                                        &rows, &nnz);
    if (rc != EOLC_OK && rc != EOLC_ERR_OVERFLOW) throw Error(rc, "collision_constraints failed");
    q.Aineq.n = rows;
    q.Aineq.rowptr.assign(rows + 1, 0);
    q.Aineq.col.resize(nnz);
    q.Aineq.val.resize(nnz);
    rc = eolc_collision_constraints(nullptr, cls.data(), (int)cls.size(), N, e, q.Aineq.rowptr.data(),
                                    q.Aineq.col.data(), q.Aineq.val.data(), rows, nnz, &rows, &nnz);
    if (rc != EOLC_OK) throw Error(rc, "collision_constraints failed");
    q.bineq.assign(rows, 0.0);
    q.Aeq.n = 3 * (int)fixedNodes.size();
    q.Aeq.rowptr.assign(q.Aeq.n + 1, 0);
    for (size_t i = 0; i < fixedNodes.size(); ++i)
        for (int k = 0; k < 3; ++k) {
            const int r = 3 * (int)i + k, dof = 3 * fixedNodes[i] + k;
            q.Aeq.col.push_back(dof);
            q.Aeq.val.push_back(1.0);
            q.Aeq.rowptr[r + 1] = r + 1;
            q.beq.push_back((1 - 0.01) * v[dof] + (scripted.empty() ? 0.0 : scripted[3 * i + k]));
        }
    return q;
}

// FixedList (FixedList.h) as parseParams fills it (parseParams.cpp:214-250, 295-304):
// from time `when` on, corner k (node ci[k]) is unused when c[k][0] == -1 ("nothing");
// otherwise axis a is fixed when c[k][a] == 1, with scripted velocity c[k][a + 3].
struct FixedList {
    double when = 0.0;
    int ci[4] = {0, 0, 0, 0};
    double c[4][6] = {{-1, -1, -1, -1, -1, -1}, {-1, -1, -1, -1, -1, -1}, {-1, -1, -1, -1, -1, -1},
                      {-1, -1, -1, -1, -1, -1}};
};

// The single-DOF equality rows of a FixedList in Constraints::fill's order
// (Constraints.cpp:472-496: corners 1..4, axes x, y, z): DOF and scripted velocity.
inline void fixedRows(const FixedList &fs, std::vector<int> &dofs, std::vector<double> &scripted)
{
    dofs.clear();
    scripted.clear();
    for (int k = 0; k < 4; ++k) {
        if (fs.c[k][0] == -1) continue;
        for (int a = 0; a < 3; ++a)
            if (fs.c[k][a] == 1.0) {
                dofs.push_back(3 * fs.ci[k] + a);
                scripted.push_back(fs.c[k][a + 3]);
            }
    }
}

// Box (Box.h): frame E1 (4x4, column-major), dimensions, velocity (angular, linear)
// as Box::v holds it.  step(h) is Box::step (Box.cpp:103-114).
struct Box {
    double dim[3] = {1.0, 1.0, 1.0};
    double E[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    double v[6] = {0, 0, 0, 0, 0, 0};
    void step(double h)
    {
        const int rc = eolc_box_step(E, v, h, E);
        if (rc != EOLC_OK) throw Error(rc, "Box::step: non-finite frame or velocity");
    }
};

// Scene::step (Scene.cpp:72-98) with remeshing and EOL off and solver "none", on the
// device-resident state (set it with Device + eolc_upload_state first):
//   Cloth::updateFix(t)            Cloth.cpp:415-420 (at most one list switch per step)
//   CD2 -> cls                     Constraints::fill's records for the host QP
//                                  (solver "none" does not read them)
//   Cloth::step                    forces, rhs, solve with the current FixedList's
//                                  rows, x += h v (eolc_step_dofs)
//   Obstacles::step(h)             every box moves; points are static
//   t += h
class Scene {
public:
    explicit Scene(Device &d) : dev_(d) {}
    double t = 0.0, h = 5e-3;
    double grav[3] = {0.0, 0.0, -9.8};
    std::vector<FixedList> fs{FixedList{}};
    size_t fsindex = 0;
    double threshold = 5e-3;          // Obstacles::cdthreshold
    std::vector<double> pts, nrm;     // static points (3P each)
    std::vector<Box> boxes;
    bool collisions = true;
    int mode = EOLC_SOLVE_FP64_BJ;
    double tol = 1e-8;
    int maxit = 100000;
    std::vector<eolc_collision> cls;  // records of the last step's CD2
    eolc_solve_stats stats{};

    void updateFix()
    {
        if (fsindex + 1 >= fs.size()) return;
        if (t > fs[fsindex + 1].when) fsindex++;
    }

    void detect()
    {
        cls.clear();
        if (pts.empty() && boxes.empty()) return;
        Obstacles o;
        o.threshold = threshold;
        o.pts = pts;
        o.nrm = nrm;
        for (const Box &b : boxes) {
            o.whd.insert(o.whd.end(), b.dim, b.dim + 3);
            o.E.insert(o.E.end(), b.E, b.E + 16);
        }
        CD2(dev_, o, cls);
    }

    void step()
    {
        updateFix();
        if (collisions) detect();
        std::vector<int> dofs;
        std::vector<double> scripted;
        fixedRows(fs[fsindex], dofs, scripted);
        dev_.check(eolc_step_dofs(dev_.ctx(), grav, h, (int)dofs.size(), dofs.data(), scripted.data(), tol, maxit,
                                  mode, &stats));
        for (Box &b : boxes) b.step(h);
        t += h;
    }

private:
    Device &dev_;
};

}  // namespace eolc
