// Drives the reference-tree subclass CLbmSolverHIP<float> (include/CLbmSolverHIPRef.hpp)
// the way CController does: the reference's CConfiguration reads an XML file,
// CDomain and CLbmSolver's constructor derive the parameters, the lid layer is
// set as in CController::setDrivenCavitySzenario (CController.cpp:1313-1345)
// and alpha / beta alternate by step parity (CController.cpp:1278-1287).
// Writes the interior velocities [3][z][y][x] and densities [z][y][x] (float32)
// to the output file.
//
//   ref_binding <config.xml> <steps> <out.bin>
//
// Built by oracle/Makefile (target `binding`) from the reference's sources
// with this repository's library; run by tests/test_reference_binding.py.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "CConfiguration.hpp"
#include "CDomain.hpp"
#include "CLbmSolverHIPRef.hpp"

// common.h declares lbm_units; the reference defines it in main.cpp.
CVector<3, int> lbm_units[] = {E0,  E1,  E2,  E3,  E4,  E5,  E6,  E7,  E8, E9,
                               E10, E11, E12, E13, E14, E15, E16, E17, E18};

int main(int argc, char** argv)
{
    if (argc != 4) {
        std::fprintf(stderr, "usage: %s config.xml steps out.bin\n", argv[0]);
        return 2;
    }
    CConfiguration<float> cfg(argv[1]);
    const int steps = std::atoi(argv[2]);
    CDomain<float> domain(0, cfg.domainSize, CVector<3, int>(0), cfg.domainLength);
    std::vector<Flag> bc(6, OBSTACLE);
    CLbmSolverHIP<float> solver(0, domain, bc, &cfg);

    CVector<3, int> size = domain.getSize();
    CVector<3, int> lidOrigin(1, size[1], 1);
    CVector<3, int> lidSize(size[0], 1, size[2]);
    std::vector<Flag> lid((size_t)size[0] * size[2], VELOCITY_INJECTION);
    solver.setFlags(lidOrigin, lidSize, lid.data());

    hipStream_t stream = NULL;   // the default stream, through the stream overloads
    for (int n = 0; n < steps; n++) {
        if (n & 1)
            solver.simulationStepBeta(&stream);
        else
            solver.simulationStepAlpha(&stream);
    }
    const size_t cells = (size_t)size[0] * size[1] * size[2];
    std::vector<float> u(3 * cells), rho(cells);
    solver.getVelocities(u.data());
    solver.getDensities(rho.data());
    std::FILE* f = std::fopen(argv[3], "wb");
    if (!f || std::fwrite(u.data(), sizeof(float), u.size(), f) != u.size() ||
        std::fwrite(rho.data(), sizeof(float), rho.size(), f) != rho.size()) {
        std::fprintf(stderr, "cannot write %s\n", argv[3]);
        return 1;
    }
    std::fclose(f);
    std::printf("ref_binding: %d steps on %dx%dx%d\n", steps, size[0], size[1], size[2]);
    return 0;
}
