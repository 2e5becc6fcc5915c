// mpsqmc_main.cpp -- command-line walk driver over the C++ host classes (the
// counterpart of the reference's main(), main.cpp:21-65, with arguments
// instead of hard-coded values).
//
//   mpsqmc_walk --L 20 --trial debug.mps --DW 2 --N 1000 --dtau 0.01 --steps 100
//               [--D 4 (random trial)] [--dmrg (trial by DMRG with bond dimension D)]
//               [--seed 1] [--fp32] [--device-rng] [--trace file] [--snake ROWSxCOLS] [--field h]
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <memory>
#include <string>

#include "mpsqmc.hpp"

using namespace mpsqmc_amd;

int main(int argc, char **argv) {
    int L = 16, D = 8, DW = -1, N = 128, steps = 10, rows = 0, cols = 0;
    unsigned seed = 20131015u;
    double dtau = 0.01, field = 0.0;
    bool fp32 = false, devrng = false, dmrg = false;
    std::string trial_file, trace;
    for (int a = 1; a < argc; ++a) {
        std::string k = argv[a];
        auto next = [&]() -> const char * {
            if (a + 1 >= argc) { std::fprintf(stderr, "missing value for %s\n", k.c_str()); std::exit(2); }
            return argv[++a];
        };
        if (k == "--L") L = std::atoi(next());
        else if (k == "--D") D = std::atoi(next());
        else if (k == "--DW") DW = std::atoi(next());
        else if (k == "--N") N = std::atoi(next());
        else if (k == "--steps") steps = std::atoi(next());
        else if (k == "--dtau") dtau = std::atof(next());
        else if (k == "--field") field = std::atof(next());
        else if (k == "--seed") seed = (unsigned)std::strtoul(next(), nullptr, 10);
        else if (k == "--trial") trial_file = next();
        else if (k == "--trace") trace = next();
        else if (k == "--fp32") fp32 = true;
        else if (k == "--device-rng") devrng = true;
        else if (k == "--dmrg") dmrg = true;
        else if (k == "--snake") {
            if (std::sscanf(next(), "%dx%d", &rows, &cols) != 2) { std::fprintf(stderr, "--snake RxC\n"); return 2; }
            L = rows * cols;
        } else { std::fprintf(stderr, "unknown argument %s\n", k.c_str()); return 2; }
    }
    try {
        Random RN(seed);
        MPSstate *trial = trial_file.empty() ? new MPSstate(L, D, 2, &RN) : new MPSstate(trial_file.c_str(), &RN);
        L = trial->gLength();
        HeisenbergMPO mpo(L, 2, false);
        if (rows > 0) {
            auto site = [&](int r, int c) { return r * cols + ((r % 2 == 0) ? c : cols - 1 - c); };
            for (int r = 0; r < rows; ++r)
                for (int c = 0; c < cols; ++c) {
                    if (c + 1 < cols) mpo.sCoupling(site(r, c), site(r, c + 1), 1.0);
                    if (r + 1 < rows) mpo.sCoupling(site(r, c), site(r + 1, c), 1.0);
                }
        } else {
            for (int i = 0; i < L - 1; ++i) mpo.sCoupling(i, i + 1, 1.0);
        }
        mpo.sField(field);
        GridGenerator grid(4);
        grid.FillMarsaglia(4);
        if (DW < 0) DW = trial->gDtrunc();
        const int prec = fp32 ? MQE_PREC_FP32 : MQE_PREC_FP64;
        const MPSQMC2::Rng rk = devrng ? MPSQMC2::Rng::Device : MPSQMC2::Rng::Host;
        std::unique_ptr<MPSQMC2> q(dmrg ? new MPSQMC2(&mpo, &grid, &RN, D, DW, N, dtau, prec, rk)
                                        : new MPSQMC2(&mpo, &grid, &RN, trial, DW, N, dtau, prec, rk));
        MPSQMC2 &qmc = *q;
        if (dmrg) std::printf("E_dmrg %.15g\n", qmc.trialEnergy());
        if (!trace.empty()) qmc.setTraceFile(trace);
        qmc.Walk(steps);
        std::printf("E_start %.15g\n", qmc.energyAtStart());
        for (const auto &r : qmc.history())
            std::printf("%d %d %.15g %.15g %.15g %.15g\n", r.step, r.nwalkers, r.projectedEnergy, r.targetEnergy,
                        r.fluctMetric, r.sumWeights);
        delete trial;
    } catch (const std::exception &ex) {
        std::fprintf(stderr, "mpsqmc_walk: %s\n", ex.what());
        return 1;
    }
    return 0;
}
