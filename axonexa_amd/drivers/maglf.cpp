// maglf — low-field saddle-coil gradient echoes in a spherical pore
// (magnetization path).
//
// The experiment of the reference driver examples/magLF_spherNeuman.cu:65-171
// (units mm, ms, T), written against axonexa_compat.hpp: walkers diffuse
// (D = 2.5e-6 mm^2/ms, dt = 0.1 us) inside a reflecting sphere of radius
// 2.5 um starting with M = (1, 0, 0); each acquisition i has two QSC_GRE
// fields (B0 = 0.5 mT, G = 0 and 6e-6 T/mm, lobes of (i + 1) 0.05 / NOA ms,
// 0.01 ms apart) whose magnetization is integrated in the lab frame
// including the concomitant gradient terms; the ADC of each acquisition is
// averaged over replicas with fresh seeds.  Seeds are explicit (i * seed,
// then seed * (r + 2) per replica; the reference uses time(NULL)).
//
//   maglf [particles=114688] [acquisitions=1000] [repeats=3] [seed=11]
// prints: diffusion time, mean ADC, ADC spread per acquisition.
#define SPECULAR_REFLECTION
#define GAMMA 267500.0
typedef double real;
#include "axonexa_compat.hpp"

#include <iomanip>

int main(int argc, char** argv) {
    const int particles = argc > 1 ? std::atoi(argv[1]) : 114688;
    const int numOfAcq = argc > 2 ? std::atoi(argv[2]) : 1000;
    const int repeats = argc > 3 ? std::atoi(argv[3]) : 3;
    const int seed = argc > 4 ? std::atoi(argv[4]) : 11;
    const real D = 2.5E-6, timestep = .0001, rad = .0025, B0 = .0005, maxDuration = 0.05;

    std::cout << std::endl << " B0 = " << B0 << std::endl << std::endl;
    magAcquisitionStream<QSC_GRE> pas(particles);
    for (int i = 0; i < numOfAcq; ++i) {
        const real gradient_duration = (i + 1) * maxDuration / numOfAcq;
        const real gradient_spacing = 0.01;
        const real echo_time = 2.0 * gradient_duration + gradient_spacing;
        const int number_of_timesteps = (int)(echo_time / timestep);
        QSC_GRE qsc1(B0, 0.0, gradient_duration, gradient_spacing);
        QSC_GRE qsc2(B0, .000006, gradient_duration, gradient_spacing);
        magAcquisition<QSC_GRE> pa(2, number_of_timesteps, i * seed);
        pa.addMeasurement(qsc1);
        pa.addMeasurement(qsc2);
        pas.addAcquisition(pa);
    }
    Sphere spher(0.0, 0.0, 0.0, rad, 0.0, 0.0, D, 0, 0.0);
    Vector3 initialM(1.0, 0.0, 0.0);
    std::vector<int> plan{0, numOfAcq, numOfAcq};
    std::vector<int> sms{14};
    for (int r = 0; r < repeats; ++r) {
        pas.runAcquisitionStream(spher, initialM, timestep, particles / 128, 128, 1, plan, sms);
        pas.calcEveryADC();
        pas.changeSeeds((long)seed * (r + 2));
    }
    std::cout << std::setprecision(20);
    for (int i = 0; i < numOfAcq; ++i) {
        auto& a = pas.getAcquisition(i);
        std::cout << a.getFieldFunctors()[0].getDiffusionTime() << " " << a.getADC() << " " << a.getADCError()
                  << std::endl;
    }
    return 0;
}
