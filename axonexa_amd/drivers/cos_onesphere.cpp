// cos_onesphere — face-centred cubic lattice of impermeable spheres, cosine
// OGSE, walkers starting outside the spheres.
//
// The experiment of the reference driver examples/Cos_OneSphere1reg.cu:
// 68-178 (units mm, ms, T), written against axonexa_compat.hpp: spheres of
// R = 1.5 um at the corners and face centres of a cubic cell sized for a
// 40 % volume fraction (14 listed spheres, four distinct under periodicity;
// the compat layer passes each once), D = 1e-6 inside / 2.5e-6 mm^2/ms
// outside, dt = 1 us, 42 ms echo; 10 acquisitions (N = 1 + 10 j periods per
// 20 ms lobe) x 30 amplitudes (G = 2.5e-6 N i T/mm) along x, all
// acquisitions in one batched launch.  Seeds are explicit (j * seed; the
// reference multiplies by time(NULL)).
//
//   cos_onesphere [particles=57344] [step_fraction=1] [seed=1234]
// prints: freq G |S| per waveform.
#define SPECULAR_REFLECTION
#define USE_INITIALIZE_IN_REGION 0
#define GAMMA 267500.0
typedef double real;
#include "axonexa_compat.hpp"

#include <iomanip>

int main(int argc, char** argv) {
    const int particles = argc > 1 ? std::atoi(argv[1]) : 57344;
    const double frac = argc > 2 ? std::atof(argv[2]) : 1.0;
    const int seed = argc > 3 ? std::atoi(argv[3]) : 1234;
    const real timestep = .001;
    const int NOI = 10, NOM = 30;

    phaseAcquisitionStream<CosGFunc> pas(particles);
    for (int j = 0; j < NOI; ++j) {
        const real gradient_duration = 20, gradient_spacing = 2.0;
        const real echo_time = 2.0 * gradient_duration + gradient_spacing;
        const int steps = std::max(2, (int)(echo_time / timestep * frac));
        phaseAcquisition<CosGFunc> pa(NOM, steps, particles, j * seed);
        for (int i = 0; i < NOM; ++i) {
            const int N = 1 + 10 * j;
            CosGFunc f(i * 0.0000025 * N, gradient_duration, gradient_spacing, N, Vector3(1.0, 0.0, 0.0));
            pa.addMeasurement(f);
        }
        pas.addAcquisition(pa);
    }

    const real radius = .0015, D_extra = 2.5E-6, D_intra = 1.0E-6, f = .4;
    const real a = std::pow((16. / 3.) * PI * radius * radius * radius / f, 1. / 3.);
    const real eps = 1e-14;
    std::vector<Sphere> basis(14);
    Lattice lattice(a, a, a, 0.0, 0.0, D_extra, 14);
    const real pos[14][3] = {{0, 0, 0}, {1, 0, 0}, {1, 0, 1}, {0, 0, 1}, {0, 1, 0}, {1, 1, 0}, {1, 1, 1},
                             {0, 1, 1}, {.5, 0, .5}, {1, .5, .5}, {.5, 1, .5}, {0, .5, .5}, {.5, .5, 1},
                             {.5, .5, 0}};
    for (int k = 0; k < 14; ++k)
        basis[k] = Sphere(pos[k][0] * a, pos[k][1] * a, pos[k][2] * a, radius, 0.0, 0.0, D_intra, k + 1, 0.0, eps);

    std::vector<int> plan{0, NOI, NOI};
    std::vector<int> sms{14};
    pas.runAcquisitionStreamLattice(&basis[0], lattice, timestep, particles / 128, 128, 1, plan, sms);

    std::cout << std::endl << " Signals " << std::endl;
    std::cout << std::setprecision(20);
    for (int i = 0; i < NOI; ++i)
        for (int j = 0; j < NOM; ++j) {
            auto& g = pas.getAcquisition(i).getGradientFunctors()[j];
            std::cout << g.getFreq() << " " << g.getG() << " " << pas.getAcquisition(i).getSignal(j) << std::endl;
        }
    return 0;
}
