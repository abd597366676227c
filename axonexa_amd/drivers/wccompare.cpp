// wccompare — streaming, WC and per-acquisition runs of one spherical pore.
//
// The experiment of the reference driver examples/WCcompare.cu:86-175 (units
// mm, ms, T), written against axonexa_compat.hpp: for NOI = 2, 4, ...
// acquisitions of two cosine-OGSE waveforms each (N = 1 + 4 j periods per
// 10 ms lobe, 2 ms gap, G = i 2.5e-6 N T/mm along x), walkers start inside
// a sphere of radius 2.5 um (D = 2.5e-6 mm^2/ms, dt = 1 us, 22 ms echo) with
// the rejection policy, and three ways of running every acquisition are
// timed:
//   runAcquisitionStream   all acquisitions in one batched launch,
//   runAcquisitionWC       Gaussian steps with rejection, one per acquisition,
//   runAcquisition         isotropic steps with rejection, one per acquisition.
// The streaming and per-acquisition runs walk the same walkers, so their
// signals agree exactly; the last line prints their largest difference.
// Seeds are explicit (j * seed; the reference multiplies by time(NULL)).
//
//   wccompare [particles=32768] [step_fraction=1] [max_noi=198] [seed=7]
// prints: NOI t_stream t_wc t_single (seconds) per NOI.
#define GAMMA 267500.0
typedef double real;
#include "axonexa_compat.hpp"

#include <iomanip>

int main(int argc, char** argv) {
    const int particles = argc > 1 ? std::atoi(argv[1]) : 32768;
    const double frac = argc > 2 ? std::atof(argv[2]) : 1.0;
    const int max_noi = argc > 3 ? std::atoi(argv[3]) : 198;
    const int seed = argc > 4 ? std::atoi(argv[4]) : 7;
    const real D = 2.5E-6, timestep = .001, duration = 10, spacing = 2.0, rad = .0025;
    const int steps = std::max(2, (int)((2.0 * duration + spacing) / timestep * frac));
    const int NOM = 2;
    double worst = 0.0;

    for (int NOI = 2; NOI <= max_noi; NOI += 2) {
        phaseAcquisitionStream<CosGFunc> pas(particles);
        for (int j = 0; j < NOI; ++j) {
            phaseAcquisition<CosGFunc> pa(NOM, steps, particles, (j + 1) * seed);
            for (int i = 0; i < NOM; ++i) {
                const int N = 1 + 4 * j;
                CosGFunc f(i * 0.0000025 * N, duration, spacing, N, Vector3(1.0, 0.0, 0.0));
                pa.addMeasurement(f);
            }
            pas.addAcquisition(pa);
        }
        Sphere spher(0.0, 0.0, 0.0, rad, 0.0, 0.0, D, 0, 0.0);
        std::vector<int> plan{0, NOI, NOI};
        std::vector<int> sms{14};

        CPUtimer t1, t2, t4;
        t1.start();
        pas.runAcquisitionStream(spher, timestep, particles / 128, 128, 1, plan, sms);
        t1.stop();
        std::vector<std::vector<real>> streamed;
        for (int j = 0; j < NOI; ++j) streamed.push_back(pas.getAcquisition(j).getMx());
        pas.flushADC();

        t2.start();
        for (int j = 0; j < NOI; ++j) pas.runAcquisitionWC(j, spher, timestep, particles / 128, 128);
        t2.stop();
        pas.flushADC();

        t4.start();
        for (int j = 0; j < NOI; ++j) pas.runAcquisition(j, spher, timestep, particles / 128, 128, 14);
        t4.stop();
        pas.flushADC();
        for (int j = 0; j < NOI; ++j)
            for (int i = 0; i < NOM; ++i)
                worst = std::max(worst, (double)std::fabs(streamed[j][i] - pas.getAcquisition(j).getMx()[i]));

        std::cout << NOI << " " << t1.getTime() << " " << t2.getTime() << " " << t4.getTime() << std::endl;
    }
    std::cout << std::setprecision(17) << "max |stream - single| = " << worst << std::endl;
    return 0;
}
