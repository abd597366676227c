// gammadist — gamma-distributed axon radii packed at random, 400 cosine-OGSE
// waveforms, independent replicas.
//
// The experiment of the reference driver examples/gammadist.cu:67-200 (units
// mm, ms, T): cylinders with radii from Gamma(k = 0.00691, rate
// 1/2.331 mm^-1) truncated to [2 sqrt(6 D dt), L/7.5], random sequential
// placement in a periodic L = 36.5 um cube; 20 frequencies (N = 1 + j
// periods per 30 ms lobe, 2 ms gap) x 20 amplitudes (G = i 2.5e-6 N T/mm);
// D = 2.5e-6 / 1e-6 mm^2/ms; dt = 1 us; 62 ms echo.  Replicas get
// independent seeds (the reference re-seeds one copy repeatedly, leaving
// six of its ten replicas identical) and the packing's periodic images come
// from the engine (no edge copies in the basis).  With this radius law
// (nearly log-uniform on the truncation interval, E[r^2] ~ 4 um^2) the
// reference's 100 cylinders would cover ~93 % of the cell, which random
// sequential placement cannot reach (the reference ignores the failed
// return of uniformCenterDist and runs on overlapping cylinders); the port
// places 50.
//
//   gammadist [particles=114688] [step_fraction=1] [replicas=10] [cylinders=50]
// prints: packing fraction, then per waveform freq G and the replica-mean
// Mx My.
#define SPECULAR_REFLECTION
#define GAMMA 267500.0
typedef double real;
#include "axonexa_compat.hpp"

#include <iomanip>

int main(int argc, char** argv) {
    const int particles = argc > 1 ? std::atoi(argv[1]) : 114688;
    const double frac = argc > 2 ? std::atof(argv[2]) : 1.0;
    const int replicas = argc > 3 ? std::atoi(argv[3]) : 10;
    const real timestep = .001;
    const int NOI = 20, NOM = 20;
    const real duration = 30, spacing = 2.0;
    const int steps = std::max(2, (int)((2.0 * duration + spacing) / timestep * frac));

    phaseAcquisitionStream<CosGFunc> pas(particles);
    phaseAcquisition<CosGFunc> pa(NOM * NOI, steps, particles, 1234);
    for (int j = 0; j < NOI; ++j)
        for (int i = 0; i < NOM; ++i) {
            const int nper = 1 + j;
            CosGFunc f(i * 0.0000025 * nper, duration, spacing, nper, Vector3(1.0, 0.0, 0.0));
            pa.addMeasurement(f);
        }
    pas.addAcquisition(pa);

    const real D_extra = 2.5E-6, D_intra = 1.0E-6, T2_i = 200, T2_e = 200;
    const real cube = .0365, alpha = .00691, beta = 1.0 / 2.331;
    const real radmin = 2.0 * std::sqrt(6.0 * D_extra * timestep);
    const int n_cyl = argc > 4 ? std::atoi(argv[4]) : 50;
    Lattice lattice(cube, cube, cube, T2_e, 0.0, D_extra, n_cyl);
    RPSLatticeInitializer<Cylinder_XY> rpsli(lattice, 0);
    rpsli.gammaRadialDist(624124, alpha, beta, radmin, cube / 7.5);
    if (!rpsli.uniformCenterDist(1208993908, 500)) {
        std::cerr << "packing failed" << std::endl;
        return 1;
    }
    rpsli.setRegions();
    rpsli.correctEdges();
    lattice = rpsli.lat;
    std::vector<Cylinder_XY> basis;
    for (size_t i = 0; i < rpsli.basis.size(); ++i) {
        basis.push_back(Cylinder_XY(0.0, 0.0, 0.0, T2_i, 0.0, D_intra, (int)i + 1, 0.0));
        basis[i].setRadius(rpsli.basis[i].getRadius());
        basis[i].setCenter(rpsli.basis[i].getCenter());
        basis[i].setRegion(rpsli.basis[i].getRegion());
    }
    std::cout << " Packing Fraction = " << rpsli.getVI() << std::endl;

    const int M = NOI * NOM;
    std::vector<double> mx(M, 0.0), my(M, 0.0);
    for (int rep = 0; rep < replicas; ++rep) {
        pas.getAcquisition(0).getSeed() = 1234 + 7919 * rep;
        pas.runAcquisitionLattice(0, &basis[0], lattice, timestep, particles / 128, 128, 14);
        for (int j = 0; j < M; ++j) {
            mx[j] += pas.getAcquisition(0).getMx()[j] / replicas;
            my[j] += pas.getAcquisition(0).getMy()[j] / replicas;
        }
    }
    std::cout << std::endl << " Signals " << std::endl;
    std::cout << std::setprecision(20);
    for (int j = 0; j < M; ++j) {
        auto& g = pas.getAcquisition(0).getGradientFunctors()[j];
        std::cout << g.getFreq() << " " << g.getG() << " " << mx[j] << " " << my[j] << " " << std::endl;
    }
    return 0;
}
