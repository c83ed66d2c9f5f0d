// Exercises the C++ facade the way a reference caller would: build bodies,
// pick a solver by id, set theta / leaf size / rule / time step through the
// reference-style accessors, call updateAccelerations() and step().
//
// usage: facade_check <solver-id> <n> <out.bin>
// Writes to out.bin: n, then x y z m (initial state), then ax ay az after
// updateAccelerations(), then x y z vx vy vz after two step() calls, all
// float32 except n (int64).  Exit status 3 when no GPU is present.
#include <cstdio>
#include <cstdlib>
#include <string>

#include "nbody_gpu.hpp"

int main(int argc, char **argv) {
    if (argc != 4) {
        std::fprintf(stderr, "usage: %s <solver-id> <n> <out.bin>\n", argv[0]);
        return 2;
    }
    const std::string id = argv[1];
    const int64_t n = std::atoll(argv[2]);
    nbody_gpu::Bodies b;
    b.resize((std::size_t)n);
    nbody_gpu::check(nbg_generate_plummer(7, n, 1.0, 1.0, 1.0, 10.0, b.x.data(), b.y.data(), b.z.data(), b.m.data(),
                                          b.vx.data(), b.vy.data(), b.vz.data()),
                     "plummer");
    if (nbody_gpu::Device::count() == 0) {
        std::fprintf(stderr, "no GPU\n");
        return 3;
    }
    try {
        nbody_gpu::Gravity rule(1.0f, 1e-4f);
        nbody_gpu::Solver::timeStep() = 1e-3f;
        auto solver = nbody_gpu::select(id, b, rule);
        auto *fmm = dynamic_cast<nbody_gpu::GpuImplicitFMMSolver *>(solver.get());
        if (!fmm) {
            solver->theta() = 0.5f;
            solver->sumMode() = NBG_SUM_ORDERED;
        }
        if (auto *bh = dynamic_cast<nbody_gpu::GpuBarnesHutSolver *>(solver.get())) bh->maxLeafSize() = 16;
        if (auto *lb = dynamic_cast<nbody_gpu::GpuLinearBVHSolver *>(solver.get())) lb->maxLeafSize() = 16;
        if (auto *qb = dynamic_cast<nbody_gpu::GpuQuadrupoleBarnesHutSolver *>(solver.get())) qb->maxLeafSize() = 16;
        if (auto *ql = dynamic_cast<nbody_gpu::GpuQuadrupoleLinearBVHSolver *>(solver.get())) ql->maxLeafSize() = 16;
        if (fmm) fmm->maxLeafSize() = 16;
        std::FILE *f = std::fopen(argv[3], "wb");
        if (!f) return 2;
        std::fwrite(&n, sizeof n, 1, f);
        for (auto *v : {&b.x, &b.y, &b.z, &b.m}) std::fwrite(v->data(), 4, (size_t)n, f);
        solver->updateAccelerations();
        for (auto *v : {&b.ax, &b.ay, &b.az}) std::fwrite(v->data(), 4, (size_t)n, f);
        solver->step();
        solver->step();
        for (auto *v : {&b.x, &b.y, &b.z, &b.vx, &b.vy, &b.vz}) std::fwrite(v->data(), 4, (size_t)n, f);
        std::fclose(f);
        std::printf("%s (%s): ok\n", solver->id().c_str(), solver->name().c_str());
    } catch (const nbody_gpu::Error &e) {
        std::fprintf(stderr, "error %d: %s\n", e.code(), e.what());
        return 1;
    }
    return 0;
}
