// cos_onecyl — one cylinder per periodic cell, 400 cosine-OGSE waveforms.
//
// The experiment of the reference driver examples/Cos_OneCyl.cu:67-165
// (units mm, ms, T), written against axonexa_compat.hpp instead of the CUDA
// headers: 20 frequencies (N = 1 + 10 j periods per 20 ms lobe, 2 ms gap)
// x 20 amplitudes (G = 8 i 2.5e-6 N T/mm) along x; cylinder R = 6 um at the
// centre of a square cell with 75 % intra-axonal fraction; D = 2.5e-6 /
// 1e-6 mm^2/ms outside / inside; dt = 1 us; 42 ms echo.  The seed is
// explicit (the reference seeds with time(NULL)) and any particle count
// works (the reference needs a multiple of the block size).  The reference's
// fraction f = 0.8 exceeds the square-packing limit pi/4, which makes its
// cylinder overlap its own periodic images (a geometry this engine rejects);
// the port uses f = 0.75.
//
//   cos_onecyl [particles=57344] [step_fraction=1] [seed=1234]
// prints: freq G Mx My per waveform.
#define SPECULAR_REFLECTION
#define GAMMA 267500.0
typedef double real;
#include "axonexa_compat.hpp"

#include <iomanip>

int main(int argc, char** argv) {
    const int particles = argc > 1 ? std::atoi(argv[1]) : 57344;
    const double frac = argc > 2 ? std::atof(argv[2]) : 1.0;
    const int seed = argc > 3 ? std::atoi(argv[3]) : 1234;
    const real timestep = .001;
    const int NOI = 20, NOM = 20;
    const real duration = 20, spacing = 2.0;
    const int steps = std::max(2, (int)((2.0 * duration + spacing) / timestep * frac));

    phaseAcquisitionStream<CosGFunc> pas(particles);
    phaseAcquisition<CosGFunc> pa(NOM * NOI, steps, particles, seed);
    for (int j = 0; j < NOI; ++j)
        for (int i = 0; i < NOM; ++i) {
            const int nper = 1 + 10 * j;
            CosGFunc f(8. * i * 0.0000025 * nper, duration, spacing, nper, Vector3(1.0, 0.0, 0.0));
            pa.addMeasurement(f);
        }
    pas.addAcquisition(pa);

    const real radius = .006, D_extra = 2.5E-6, D_intra = 1.0E-6, T2_i = 200, T2_e = 200, f = .75;
    const real a = std::sqrt(PI * radius * radius / f);
    std::vector<Cylinder_XY> basis(1);
    Lattice lattice(a, a, a, T2_e, 0.0, D_extra, 1);
    basis[0] = Cylinder_XY(a / 2.0, a / 2.0, radius, T2_i, 0.0, D_intra, 1, 0.0);

    pas.runAcquisitionLattice(0, &basis[0], lattice, timestep, particles / 128, 128, 14);

    std::cout << std::endl << " Signals " << std::endl;
    std::cout << std::setprecision(20);
    for (int j = 0; j < NOI * NOM; ++j) {
        auto& g = pas.getAcquisition(0).getGradientFunctors()[j];
        std::cout << g.getFreq() << " " << g.getG() << " " << pas.getAcquisition(0).getMx()[j] << " "
                  << pas.getAcquisition(0).getMy()[j] << " " << std::endl;
    }
    return 0;
}
