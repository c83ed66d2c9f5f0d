// nbody_gpu.hpp — C++ host facade over the nbg C ABI (include/nbg.h).
//
// It mirrors the reference simulator's solver interface for the tree-code
// force pass (JacksonCampolattaro/n-body, paths relative to src/NBody/):
//   Solver<Rule>             Simulation/Solver.h:26-187     step(), updateAccelerations(),
//                                                          timeStep(), id(), name()
//   ActiveTreeSolver         Solvers/ActiveTreeSolver.h:17-80   theta()
//   BarnesHutSolver          Solvers/BarnesHutSolver.h:14-32    maxDepth(), maxLeafSize(), "barnes-hut"
//   LinearBVHSolver          Solvers/LinearBVHSolver.h:18-28    maxLeafSize(), "linear-bvh"
//   Gravity(G, epsilon)      Physics/Rules/Gravity.h:185-209    epsilon is added to r^2
//   MultiSolver::select(id)  Simulation/Solvers/MultiSolver.h:79-87
//
// Like the reference, a solver holds non-owning references to the simulation
// (here `Bodies`, SoA float32 host arrays standing in for the EnTT registry's
// Position / Velocity / Acceleration / Mass components) and to the rule, and
// parameter accessors return references that take effect at the next force
// pass.  step() has the reference's semantics: afterwards the host positions
// and velocities hold the new state.  run(n) keeps the state on the device
// for n steps and copies it back once.  Errors (which the reference never
// signals) are thrown as nbody_gpu::Error with nbg_last_error()'s message.
#pragma once

#include <array>
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/nbg.h"

namespace nbody_gpu {

class Error : public std::runtime_error {
public:
    Error(int code, const std::string &what) : std::runtime_error(what), code_(code) {}
    int code() const { return code_; }

private:
    int code_;
};

inline void check(int rc, const char *what) {
    if (rc < 0) throw Error(rc, std::string(what) + ": " + nbg_last_error());
}

// Gravity(G, epsilon) (Physics/Rules/Gravity.h:185-195): epsilon is added to
// r^2, i.e. it is the square of a softening length.
struct Gravity {
    float G = 1.0f;
    float epsilon = 0.00001f;
    Gravity() = default;
    Gravity(float g, float eps) : G(g), epsilon(eps) {}
    float &g() { return G; }
    float &eps() { return epsilon; }
};

// Body state: the Position, Velocity, Acceleration and Mass components.
struct Bodies {
    std::vector<float> x, y, z, m, vx, vy, vz, ax, ay, az;
    std::size_t size() const { return x.size(); }
    void resize(std::size_t n) {
        for (auto *v : {&x, &y, &z, &m, &vx, &vy, &vz, &ax, &ay, &az}) v->assign(n, 0.0f);
    }
};

// One GPU (nbg_ctx).
class Device {
public:
    explicit Device(int index = 0) { check(nbg_ctx_create(index, &ctx_), "nbg_ctx_create"); }
    ~Device() { nbg_ctx_destroy(ctx_); }
    Device(const Device &) = delete;
    Device &operator=(const Device &) = delete;
    nbg_ctx *handle() const { return ctx_; }
    static int count() { return nbg_device_count(); }

private:
    nbg_ctx *ctx_ = nullptr;
};

// Solver<Gravity> (Simulation/Solver.h:26-187).
class Solver {
public:
    virtual ~Solver() { nbg_sys_destroy(sys_); }
    Solver(const Solver &) = delete;
    Solver &operator=(const Solver &) = delete;

    virtual std::string id() = 0;
    virtual std::string name() = 0;

    // Solver::timeStep() (Solver.h:101): one step size shared by every solver.
    static float &timeStep() {
        static float dt = 0.001f;
        return dt;
    }

    // Summation order on the device: ORDERED (bit-exact with the reference's
    // arithmetic) or FAST (same interaction set, rsqrt + FMA).
    int &sumMode() { return prm_.sum_mode; }

    // Solver::updateAccelerations (ActiveTreeSolver.h:33-72): positions and
    // masses in, accelerations out (host arrays).
    virtual void updateAccelerations() {
        upload();
        check(nbg_sys_update_accelerations(sys_, sim_.ax.data(), sim_.ay.data(), sim_.az.data()),
              "updateAccelerations");
    }

    // Solver::step (Solver.h:103-113): updateAccelerations, v += a*dt, p += v*dt.
    void step() { run(1); }

    // nsteps device-resident steps, then one copy of the state back.
    void run(int nsteps) {
        upload();
        check(nbg_sys_step(sys_, nsteps), "step");
        check(nbg_sys_synchronize(sys_), "step");
        check(nbg_sys_get_state(sys_, sim_.x.data(), sim_.y.data(), sim_.z.data(), sim_.vx.data(), sim_.vy.data(),
                                sim_.vz.data()),
              "get_state");
        if (nsteps > 0)
            check(nbg_sys_get_accelerations(sys_, sim_.ax.data(), sim_.ay.data(), sim_.az.data()), "accelerations");
    }

    // Interaction counts of the last force pass (SimpleTrackingRule, 64-bit).
    std::pair<uint64_t, uint64_t> interactionCounts() {
        uint64_t pp = 0, np = 0;
        check(nbg_sys_counts(sys_, &pp, &np), "counts");
        return {pp, np};
    }

    // All four counters {pp, np, pn, nn} (pn and nn: the implicit FMM's
    // particle-node and node-node calls).
    std::array<uint64_t, 4> interactionCounts4() {
        std::array<uint64_t, 4> c{};
        check(nbg_sys_counts4(sys_, c.data()), "counts4");
        return c;
    }

    Bodies &simulation() { return sim_; }
    Gravity &rule() { return rule_; }
    nbg_sys *handle() { return sys_; }

protected:
    // theta / sum_mode: the solver's own defaults where they differ from
    // nbg_params_default's (negative: keep those).
    Solver(Bodies &sim, Gravity &rule, int solver_kind, const Device &dev = default_device(), float theta = -1.0f,
           int sum_mode = -1)
        : sim_(sim), rule_(rule), dev_(dev) {
        nbg_params_default(&prm_);
        prm_.solver = solver_kind;
        if (theta >= 0.0f) prm_.theta = theta;
        if (sum_mode >= 0) prm_.sum_mode = sum_mode;
        if (sim_.ax.size() != sim_.size()) {
            sim_.ax.assign(sim_.size(), 0.0f);
            sim_.ay.assign(sim_.size(), 0.0f);
            sim_.az.assign(sim_.size(), 0.0f);
        }
        apply_rule();
        n_ = (int64_t)sim_.size();
        check(nbg_sys_create(dev_.handle(), &prm_, n_, &sys_), "nbg_sys_create");
        applied_ = prm_;
    }

    static const Device &default_device() {
        static Device d(0);
        return d;
    }

    nbg_params prm_{};

private:
    void apply_rule() {
        prm_.G = rule_.G;
        prm_.eps2 = rule_.epsilon;
        prm_.dt = timeStep();
    }

    // Pushes parameters changed through the accessors and the host state.
    void upload() {
        if ((int64_t)sim_.size() != n_ || (int64_t)sim_.ax.size() != n_)
            throw Error(NBG_ERR_ARG, "body count changed since the solver was created");
        apply_rule();
        if (std::memcmp(&prm_, &applied_, sizeof prm_) != 0) {
            check(nbg_sys_set_params(sys_, &prm_), "set_params");
            applied_ = prm_;
        }
        check(nbg_sys_set_state(sys_, sim_.x.data(), sim_.y.data(), sim_.z.data(), sim_.m.data(), sim_.vx.data(),
                                sim_.vy.data(), sim_.vz.data()),
              "set_state");
    }
    Bodies &sim_;
    Gravity &rule_;
    const Device &dev_;
    nbg_sys *sys_ = nullptr;
    int64_t n_ = 0;
    nbg_params applied_{};
};

// ActiveTreeSolver (Solvers/ActiveTreeSolver.h:17-80).
class ActiveTreeSolver : public Solver {
public:
    float &theta() { return prm_.theta; }

protected:
    using Solver::Solver;
};

// BarnesHutSolver (Solvers/BarnesHutSolver.h:14-32): octree, SideLengthOverDistance.
class GpuBarnesHutSolver : public ActiveTreeSolver {
public:
    GpuBarnesHutSolver(Bodies &sim, Gravity &rule, const Device &dev = default_device())
        : ActiveTreeSolver(sim, rule, NBG_BARNES_HUT, dev) {}
    std::string id() override { return "gpu-barnes-hut"; }
    std::string name() override { return "Barnes-Hut (MI355X)"; }
    int &maxDepth() { return prm_.max_depth; }
    int &maxLeafSize() { return prm_.leaf_size; }
};

// LinearBVHSolver (Solvers/LinearBVHSolver.h:18-28): Morton-ordered BVH,
// DiagonalOverCenterDistance.
class GpuLinearBVHSolver : public ActiveTreeSolver {
public:
    GpuLinearBVHSolver(Bodies &sim, Gravity &rule, const Device &dev = default_device())
        : ActiveTreeSolver(sim, rule, NBG_LINEAR_BVH, dev) {}
    std::string id() override { return "gpu-linear-bvh"; }
    std::string name() override { return "Linear-BVH (MI355X)"; }
    int &maxLeafSize() { return prm_.leaf_size; }
    // MortonCode width (Solvers/MortonSort.h:25): 63 (default) or 30 bits.
    int &mortonBits() { return prm_.morton_bits; }
};

// QuadrupoleBarnesHutSolver = MultipoleBarnesHutSolver<2>
// (Solvers/BarnesHutSolver.h:34-66): the octree with second moments.
class GpuQuadrupoleBarnesHutSolver : public ActiveTreeSolver {
public:
    GpuQuadrupoleBarnesHutSolver(Bodies &sim, Gravity &rule, const Device &dev = default_device())
        : ActiveTreeSolver(sim, rule, NBG_BARNES_HUT_QUADRUPOLE, dev) {}
    std::string id() override { return "gpu-barnes-hut-4p"; }
    std::string name() override { return "Barnes-Hut (Quadrupole, MI355X)"; }
    int &maxDepth() { return prm_.max_depth; }
    int &maxLeafSize() { return prm_.leaf_size; }
};

// QuadrupoleLinearBVHSolver = MultipoleLinearBVHSolver<2>
// (Solvers/LinearBVHSolver.h:30-53): the LBVH with second moments.
class GpuQuadrupoleLinearBVHSolver : public ActiveTreeSolver {
public:
    GpuQuadrupoleLinearBVHSolver(Bodies &sim, Gravity &rule, const Device &dev = default_device())
        : ActiveTreeSolver(sim, rule, NBG_LINEAR_BVH_QUADRUPOLE, dev) {}
    std::string id() override { return "gpu-linear-bvh-4p"; }
    std::string name() override { return "Linear BVH (Quadrupole, MI355X)"; }
    int &maxLeafSize() { return prm_.leaf_size; }
    int &mortonBits() { return prm_.morton_bits; }
};

// ImplicitFMMSolver = MultipoleImplicitFMMSolver<1> (Solvers/ImplicitFMMSolver.h:14-36):
// the implicit dual traversal over the octree, theta 0.4
// (ImplicitDualTraversalSolver.h:22); FAST sum mode only.
class GpuImplicitFMMSolver : public ActiveTreeSolver {
public:
    GpuImplicitFMMSolver(Bodies &sim, Gravity &rule, const Device &dev = default_device())
        : ActiveTreeSolver(sim, rule, NBG_IMPLICIT_FMM, dev, 0.4f, NBG_SUM_FAST) {}
    std::string id() override { return "gpu-implicit-fmm-2p"; }
    std::string name() override { return "Implicit Fast Multipole Method (1-pole, MI355X)"; }
    int &maxDepth() { return prm_.max_depth; }
    int &maxLeafSize() { return prm_.leaf_size; }
};

// NaiveSolver (Solvers/NaiveSolver.h:14-52): direct summation, no tree.
class GpuNaiveSolver : public Solver {
public:
    GpuNaiveSolver(Bodies &sim, Gravity &rule, const Device &dev = default_device())
        : Solver(sim, rule, NBG_NAIVE, dev) {}
    std::string id() override { return "gpu-naive"; }
    std::string name() override { return "Naive (MI355X)"; }
};

// MultiSolver::select(id) (Simulation/Solvers/MultiSolver.h:79-87) for the
// device tree solvers.
inline std::unique_ptr<ActiveTreeSolver> select(const std::string &id, Bodies &sim, Gravity &rule) {
    if (id == "gpu-barnes-hut") return std::make_unique<GpuBarnesHutSolver>(sim, rule);
    if (id == "gpu-linear-bvh") return std::make_unique<GpuLinearBVHSolver>(sim, rule);
    if (id == "gpu-barnes-hut-4p") return std::make_unique<GpuQuadrupoleBarnesHutSolver>(sim, rule);
    if (id == "gpu-linear-bvh-4p") return std::make_unique<GpuQuadrupoleLinearBVHSolver>(sim, rule);
    if (id == "gpu-implicit-fmm-2p") return std::make_unique<GpuImplicitFMMSolver>(sim, rule);
    throw Error(NBG_ERR_ARG, "unknown solver id " + id);
}

}  // namespace nbody_gpu
