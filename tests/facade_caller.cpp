// A C++ caller of the DEM facade (include/xDiscreteElementMethodSimulation.hpp), written the
// way the reference's simulator drives xIntegratorVV (xDynamicSimulator.cpp:518-537): build a
// scene, Initialize, OneStepSimulation per step, SaveStepResult. Writes the host copies of
// position and velocity (scene 0, reference layouts) and the device error word to a file.
//   facade_caller <workload> <seed> <np_override> <nsteps> <out.bin>
#include <cstdio>
#include <cstdlib>

#include "xDiscreteElementMethodSimulation.hpp"

int main(int argc, char** argv)
{
    if (argc != 6) {
        std::fprintf(stderr, "usage: %s workload seed np_override nsteps out\n", argv[0]);
        return 2;
    }
    const int workload = std::atoi(argv[1]);
    const unsigned long long seed = std::strtoull(argv[2], nullptr, 10);
    const unsigned npo = (unsigned)std::strtoul(argv[3], nullptr, 10);
    const unsigned nsteps = (unsigned)std::strtoul(argv[4], nullptr, 10);
    xdem_scene scene;
    if (xdem_scene_build(workload, seed, npo, 0, &scene) != XDEM_OK) return 3;
    xIntegratorVV vv;
    int rc = vv.Initialize(&scene, 1, nullptr);   // default configuration: FP64, parity quirks
    if (rc != XDEM_OK) return 4;
    double ct = 0.0;
    for (unsigned k = 1; k <= nsteps; k++) {
        ct += scene.dt;
        if ((rc = vv.OneStepSimulation(ct, k)) != XDEM_OK) return 5;
    }
    if (!vv.SaveStepResult(nsteps)) return 6;
    const unsigned err = vv.ErrorWord();
    std::FILE* f = std::fopen(argv[5], "wb");
    if (!f) return 7;
    const unsigned n = vv.num_particles();
    std::fwrite(&n, sizeof(n), 1, f);
    std::fwrite(&err, sizeof(err), 1, f);
    std::fwrite(vv.HostPosition().data(), sizeof(double), vv.HostPosition().size(), f);
    std::fwrite(vv.HostVelocity().data(), sizeof(double), vv.HostVelocity().size(), f);
    std::fclose(f);
    xdem_scene_free(&scene);
    std::printf("facade_caller: %u spheres, %u steps, error word %u\n", n, nsteps, err);
    return 0;
}
