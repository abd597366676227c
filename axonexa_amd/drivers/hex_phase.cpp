// hex_phase — hexagonal lattice of impermeable cylinders at four packing
// fractions, cosine OGSE across the cylinders.
//
// The experiment of the reference driver examples/hexPhase.cu:58-136 (units
// mm, ms, T), written against axonexa_compat.hpp: cylinders R = 2.5 um in the
// rectangular cell (d, d sqrt 3) with d = sqrt(2 pi R^2 / (sqrt 3 f)),
// f = 0.6, 0.7, 0.8, 0.9; four listed cylinders, two distinct under
// periodicity (the compat layer passes each once); D = 1e-6 inside /
// 2.5e-6 mm^2/ms outside, T2 200 ms (not applied: no USE_RELAXATION),
// dt = 1 us, 22 ms echo; 3 acquisitions (N = 1 + j periods per 10 ms lobe)
// x 2 amplitudes (G = 0 and 2.5e-6 N T/mm) along x.  Seeds are explicit
// (j * seed; the reference uses time(NULL)).
//
//   hex_phase [particles=14336] [step_fraction=1] [seed=1234]
// prints: "f = ..." then freq G |S| per waveform, per packing fraction.
#define SPECULAR_REFLECTION
#define GAMMA 267500.0
typedef double real;
#include "axonexa_compat.hpp"

#include <iomanip>

int main(int argc, char** argv) {
    const int particles = argc > 1 ? std::atoi(argv[1]) : 114688 / 8;
    const double frac = argc > 2 ? std::atof(argv[2]) : 1.0;
    const int seed = argc > 3 ? std::atoi(argv[3]) : 1234;
    const real timestep = .001;
    const int NOI = 3, NOM = 2;

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
    const real s3 = 0.86602540378443864676372317075294;
    for (int r = 0; r < 4; ++r) {
        const real radius = .0025, D_extra = 2.5E-6, D_intra = 1.0E-6, T2_i = 200, T2_e = 200;
        const real f = .6 + r * .1;
        std::cout << "f = " << f << std::endl;
        const real d = std::sqrt(2.0 * PI * radius * radius / (std::sqrt(3.0) * f));
        Cylinder_XY cyls[4];
        Lattice lattice(d, 2.0 * d * s3, d, T2_e, 0.0, D_extra, 4);
        cyls[0] = Cylinder_XY(d / 2.0, 0.0, radius, T2_i, 0.0, D_intra, 1, 0.0);
        cyls[1] = Cylinder_XY(0.0, d * s3, radius, T2_i, 0.0, D_intra, 2, 0.0);
        cyls[2] = Cylinder_XY(d / 2.0, 2.0 * d * s3, radius, T2_i, 0.0, D_intra, 3, 0.0);
        cyls[3] = Cylinder_XY(d, d * s3, radius, T2_i, 0.0, D_intra, 4, 0.0);
        for (int i = 0; i < NOI; ++i) pas.runAcquisitionLattice(i, cyls, lattice, timestep, particles / 128, 128, 1);
        std::cout << std::setprecision(20);
        for (int i = 0; i < NOI; ++i)
            for (int j = 0; j < NOM; ++j) {
                auto& g = pas.getAcquisition(i).getGradientFunctors()[j];
                std::cout << g.getFreq() << " " << g.getG() << " " << pas.getAcquisition(i).getSignal(j) << std::endl;
            }
    }
    return 0;
}
