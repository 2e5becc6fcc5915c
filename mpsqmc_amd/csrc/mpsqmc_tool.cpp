// mpsqmc_tool.cpp -- runs one MPSstate / Walker method of the C++ host classes
// on an .mps file (the reference's text format, MPSstate.cpp:92-128,661-686) and
// writes the result, so that each method can be checked against the FP64
// oracle and dense vectors.  Host code only: no GPU is used.
//
//   mpsqmc_tool <method> --in a.mps [--out b.mps] [--snake RxC] [--field h] [--dtau t] ...
//     svd-term     --first i --second j --a a --b b [--hc]  ApplyTwoSiteTrotterTerm (SVD indices)
//     grid-term    --first i --second j --point g           ApplyTwoSiteTrotterTerm (grid point)
//     field                                                  ApplyOneSiteTrotterTermEverywhere
//     mpo                                                    ApplyMPO; prints <in|H in>
//     right-normalize                                        RightNormalize; prints the norm
//     scalar       --factor f                                ScalarMultiplication
//     walker       --trial t.mps                             Walker::setOverlap and
//                                                            calcIndividualProjectedEnergy(H trial);
//                                                            prints O and E_L
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>

#include "mpsqmc.hpp"

using namespace mpsqmc_amd;

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: mpsqmc_tool <method> --in a.mps [options]\n");
        return 2;
    }
    const std::string method = argv[1];
    std::string in, out, trial;
    int rows = 0, cols = 0, first = 0, second = 1, a = 0, b = 0, point = 0;
    double field = 0.0, dtau = 0.01, factor = 1.0;
    bool hc = false;
    for (int q = 2; q < argc; ++q) {
        const std::string k = argv[q];
        auto next = [&]() -> const char * {
            if (q + 1 >= argc) { std::fprintf(stderr, "missing value for %s\n", k.c_str()); std::exit(2); }
            return argv[++q];
        };
        if (k == "--in") in = next();
        else if (k == "--out") out = next();
        else if (k == "--trial") trial = next();
        else if (k == "--first") first = std::atoi(next());
        else if (k == "--second") second = std::atoi(next());
        else if (k == "--a") a = std::atoi(next());
        else if (k == "--b") b = std::atoi(next());
        else if (k == "--point") point = std::atoi(next());
        else if (k == "--field") field = std::atof(next());
        else if (k == "--dtau") dtau = std::atof(next());
        else if (k == "--factor") factor = std::atof(next());
        else if (k == "--hc") hc = true;
        else if (k == "--snake") {
            if (std::sscanf(next(), "%dx%d", &rows, &cols) != 2) { std::fprintf(stderr, "--snake RxC\n"); return 2; }
        } else { std::fprintf(stderr, "unknown argument %s\n", k.c_str()); return 2; }
    }
    try {
        if (in.empty()) throw std::invalid_argument("--in is required");
        MPSstate st(in.c_str(), nullptr);
        const int L = st.gLength();
        HeisenbergMPO mpo(L, st.gPhys_d(), false);
        if (rows > 0) {
            if (rows * cols != L) throw std::invalid_argument("--snake does not match the chain length");
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
        TrotterHeisenberg trotter(&mpo, dtau);
        GridGenerator grid(st.gPhys_d() * st.gPhys_d());
        grid.FillMarsaglia(4);
        MPSstate result(&st);
        if (method == "svd-term") {
            result.ApplyTwoSiteTrotterTerm(&trotter, first, second, a, b, hc);
        } else if (method == "grid-term") {
            result.ApplyTwoSiteTrotterTerm(&trotter, first, second, &grid, point);
        } else if (method == "field") {
            result.ApplyOneSiteTrotterTermEverywhere(&trotter);
        } else if (method == "mpo") {
            result.ApplyMPO(&mpo, &st);
            std::printf("energy %.17g\n", st.InnerProduct(&result));
        } else if (method == "right-normalize") {
            std::printf("norm %.17g\n", result.RightNormalize());
        } else if (method == "scalar") {
            result.ScalarMultiplication(factor);
        } else if (method == "walker") {
            if (trial.empty()) throw std::invalid_argument("walker needs --trial");
            MPSstate t(trial.c_str(), nullptr);
            MPSstate ht(&t);
            ht.ApplyMPO(&mpo, &t);
            Walker w(&st, NAN, 1.0, 0.0);
            w.setOverlap(&t);
            std::printf("overlap %.17g\n", w.gOverlap());
            std::printf("energy %.17g\n", w.calcIndividualProjectedEnergy(&ht));
            result = MPSstate(w.gState());
        } else {
            throw std::invalid_argument("unknown method " + method);
        }
        if (!out.empty()) {
            std::ofstream f(out);
            f << result;
            if (!f) throw std::runtime_error("cannot write " + out);
        }
    } catch (const std::exception &ex) {
        std::fprintf(stderr, "mpsqmc_tool: %s\n", ex.what());
        return 1;
    }
    return 0;
}
