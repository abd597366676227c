// plane_phase — restricted diffusion between two parallel planes, cosine
// OGSE across them, ADC against frequency.
//
// The experiment of the reference driver examples/planePhase.cu:63-152
// (units mm, ms, T), written against axonexa_compat.hpp: planes at x = 0
// (normal +x) and x = l (normal -x) in an l x l x l cell, l = 1, 2, 5, 10 um
// (the two are one plane under periodicity; the compat layer passes it
// once); D = 2.5e-6 mm^2/ms, dt = 1 us, 22 ms echo; 10 acquisitions
// (N = 1 + j periods per 10 ms lobe) x 2 amplitudes (G = 0 and 2.5e-6 N
// T/mm) along x; three replicas (runAcquisitionStreamLattice, calcEveryADC,
// changeSeeds) per length.  Seeds are explicit (j * seed, then
// seed * r * (j + 1) for replica r; the reference uses time(NULL)).
//
//   plane_phase [particles=4096] [step_fraction=1] [seed=1234]
// prints: " PLANE LENGTH = l" then freq ADC per acquisition.
#define SPECULAR_REFLECTION
#define GAMMA 267500.0
typedef double real;
#include "axonexa_compat.hpp"

#include <iomanip>

int main(int argc, char** argv) {
    const int particles = argc > 1 ? std::atoi(argv[1]) : 4096;
    const double frac = argc > 2 ? std::atof(argv[2]) : 1.0;
    const int seed = argc > 3 ? std::atoi(argv[3]) : 1234;
    const real timestep = .001;
    const int NOI = 10, NOM = 2;

    phaseAcquisitionStream<CosGFunc> pas(particles);
    for (int j = 0; j < NOI; ++j) {
        const real gradient_duration = 10, gradient_spacing = 2.0;
        const int steps = std::max(2, (int)((2.0 * gradient_duration + gradient_spacing) / timestep * frac));
        phaseAcquisition<CosGFunc> pa(NOM, steps, particles, j * seed);
        for (int i = 0; i < NOM; ++i) {
            const int N = 1 + j;
            CosGFunc f(i * 0.0000025 * N, gradient_duration, gradient_spacing, N, Vector3(1.0, 0.0, 0.0));
            pa.addMeasurement(f);
        }
        pas.addAcquisition(pa);
    }
    const real lengths[4] = {.001, .002, .005, .010};
    for (int r = 0; r < 4; ++r) {
        const real D_extra = 2.5E-6, T2_e = 200;
        const real l = lengths[r];
        std::cout << " PLANE LENGTH = " << l << std::endl;
        Plane cyls[2];
        Lattice lattice(l, l, l, T2_e, 0.0, D_extra, 2);
        cyls[0] = Plane(Vector3(0.0, 0.0, 0.0), Vector3(1.0, 0.0, 0.0));
        cyls[1] = Plane(Vector3(l, 0.0, 0.0), Vector3(-1.0, 0.0, 0.0));
        std::vector<int> plan{0, NOI, NOI};
        std::vector<int> sms{1};
        for (int i = 0; i < 3; ++i) {
            pas.runAcquisitionStreamLattice(cyls, lattice, timestep, particles / 128, 128, 1, plan, sms);
            pas.calcEveryADC();
            pas.changeSeeds((long)seed * (i + 1));
        }
        std::cout << std::setprecision(20);
        for (int i = 0; i < NOI; ++i)
            std::cout << pas.getAcquisition(i).getGradientFunctors()[0].getFreq() << " "
                      << pas.getAcquisition(i).getADC() << std::endl;
        pas.flushADC();
    }
    return 0;
}
