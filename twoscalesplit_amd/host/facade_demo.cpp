// facade_demo.cpp -- the reference's init()/display() sequence
// (TwoScale2D/main.cpp:131-208, 65-102) written against the class API kept
// by WCSPHSolver.h, headless: builds the reference's demo scene, advances
// it and prints a one-line JSON summary of the read-back state.
//
// usage: facade_demo [steps] [state-file]
// With a state file it also saves the scene there, advances 20 steps,
// reloads it, advances the same 20 steps and reports whether both runs
// ended on the same positions ("resume_equal").
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <vector>

#include "WCSPHSolver.h"

int main(int argc, char** argv) {
    const int steps = argc > 1 ? std::atoi(argv[1]) : 100;
    try {
        // R/main.cpp:144-175 (the reference's demo scene)
        const double PI = 3.14159265358979323846;
        float restDensity = 1000.0f;
        float fluidVol = 100.0f * 0.0025f * 150.0f * 0.0025f;
        unsigned int numParticles = 101 * 151;
        float particleMass = fluidVol * restDensity / numParticles;
        float effectiveRadius = (float)std::sqrt(fluidVol * 25.0f / (numParticles * PI));
        float effectiveRadiusHigh = 0.5f * effectiveRadius;
        float taitCoeff = 1119.0714f;
        float speedSound = 88.1472f;
        float alpha = 0.04f;
        float tensionCoefficient = 0.08f;
        float timeStep = 0.0002f;

        std::unique_ptr<ParticleSystem> low(
            CreateParticleBox(0.205f, 0.105f, 0.0025f, 101, 151, particleMass));
        std::unique_ptr<ParticleSystem> high(
            new ParticleSystem(low->GetMaxParticles() * 4, low->GetMass() / 4.0f));
        std::unique_ptr<ParticleSystem> bnd(CreateParticleBoxCanvas(
            0.1f, 0.1f, 0.0025f, 321, 321, 3, 3, 1000.0f * 0.25f * 0.25f / (101.0f * 151.0f)));
        WCSPHConfig config(0.0f, 0.0f, 1.0f, 1.0f, effectiveRadius, effectiveRadiusHigh,
                           restDensity, taitCoeff, speedSound, alpha, tensionCoefficient,
                           timeStep);
        WCSPHSolver solver(config, *low, *high, *bnd);
        solver.Bind();
        for (int i = 0; i < steps; ++i) solver.AdvanceTS();

        low->Map();
        high->Map();
        const unsigned nl = low->GetNumParticles(), nh = high->GetNumParticles();
        double rho_mean = 0.0, ymin = 1e30, ymax = -1e30;
        int finite = 1;
        for (unsigned k = 0; k < nl; ++k) {
            int id = low->ParticleIDs()[k];
            float y = low->Positions()[2 * id + 1];
            rho_mean += low->Densities()[id];
            ymin = std::fmin(ymin, y);
            ymax = std::fmax(ymax, y);
            finite &= std::isfinite(y) && std::isfinite(low->Densities()[id]);
        }
        rho_mean /= nl ? nl : 1;
        low->Unmap();
        high->Unmap();
        int resume_equal = -1;
        if (argc > 2) {
            auto run20 = [&]() {
                for (int i = 0; i < 20; ++i) solver.AdvanceTS();
                low->Map();
                high->Map();
                std::vector<float> p(low->Positions(),
                                     low->Positions() + 2 * low->GetMaxParticles());
                p.insert(p.end(), high->Positions(),
                         high->Positions() + 2 * high->GetMaxParticles());
                low->Unmap();
                high->Unmap();
                return p;
            };
            solver.SaveState(argv[2]);
            const std::vector<float> a = run20();
            solver.LoadState(argv[2]);
            const std::vector<float> b = run20();
            resume_equal = std::memcmp(a.data(), b.data(), a.size() * sizeof(float)) == 0;
        }
        std::printf("{\"steps\": %d, \"n_low\": %u, \"n_high\": %u, \"rho_mean\": %.6f, "
                    "\"y_min\": %.6f, \"y_max\": %.6f, \"finite\": %d, \"resume_equal\": %d}\n",
                    steps, nl, nh, rho_mean, ymin, ymax, finite, resume_equal);
        solver.Unbind();
    } catch (const std::runtime_error& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
