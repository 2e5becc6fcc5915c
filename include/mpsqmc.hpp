// mpsqmc.hpp -- C++ host classes with the reference's driver interface
// (include/MPSQMC2.h, Walker.h, MPSstate.h, TrotterHeisenberg.h,
// GridGenerator.h, HeisenbergMPO.h, Random.h), implemented over the engine's
// C ABI (mqe.h).  MPSQMC2::Walk keeps the reference's call order
// (MPSQMC2.cpp:443-513): propagate all walkers, scaling and E_T, projected
// energy and fluctuation metric, trace line, population control -- with the
// walker population resident on the GPU.
#ifndef MPSQMC_HPP
#define MPSQMC_HPP

#include <ostream>
#include <string>
#include <vector>

#include "mqe.h"

namespace mpsqmc_amd {

// Random.h: one seeded mt19937 stream, u = x * 2^-32 (the reference seeds
// from time(0), Random.cpp:31; a fixed seed makes runs reproducible).
class Random {
  public:
    explicit Random(unsigned int seed = 20131015u);
    ~Random();
    Random(const Random &) = delete;
    Random &operator=(const Random &) = delete;
    double rand();
    void fill(long n, double *out);
    void peek(long n, double *out);
    void discard(long k);
    // the seed; MPSQMC2 also keys the device random numbers with it
    unsigned int seed() const { return seed_; }

  private:
    mqe_rng *rng_;
    unsigned int seed_;
};

// GridGenerator.h:13-34
class GridGenerator {
  public:
    explicit GridGenerator(int dim);
    int gDim() const { return dim_; }
    int gNPoints() const { return n_; }
    double gCoOfPoint(int point, int index) const { return grid_[index + dim_ * point]; }
    void FillMarsaglia(int pointsPerCo);
    const double *data() const { return grid_.data(); }

  private:
    int dim_;
    int n_ = 0;
    std::vector<double> grid_;
};

// HeisenbergMPO.h:25-52 (couplings and field; H = sum J_ij S_i.S_j - h sum S^z_i)
class HeisenbergMPO {
  public:
    HeisenbergMPO(int length, int phys_d, bool useLadder = false);
    void sCoupling(int i, int j, double value);
    double gCoupling(int i, int j) const;
    void sField(double value) { field_ = value; }
    double gField() const { return field_; }
    int gLength() const { return L_; }
    int gPhys_d() const { return d_; }

  private:
    int L_, d_;
    double field_ = 0.0;
    std::vector<double> J_;  // J[(j*(j-1))/2 + i], i < j
};

// TrotterHeisenberg.h:16-40
class TrotterHeisenberg {
  public:
    TrotterHeisenberg(const HeisenbergMPO *mpo, double dtau);
    bool gIsMagneticField() const { return field_ != 0.0; }
    double gField() const { return field_; }
    double gCoupling(int i, int j) const { return mpo_->gCoupling(i, j); }
    double gTwoSitePropSVD_Sing(double J, int k) const;
    double gTwoSitePropSVD_Left(double J, int k, int i, int j) const;
    double gTwoSitePropSVD_Right(double J, int k, int i, int j) const;
    double gSingleSiteProp(int i, int j) const;
    // flattened per-coupling tables for a coupling list
    void tables(const std::vector<double> &Jc, std::vector<double> &S, std::vector<double> &U,
                std::vector<double> &VT) const;

  private:
    int find(double J) const;
    const HeisenbergMPO *mpo_;
    double dtau_, field_;
    std::vector<double> dJ_, S_, U_, VT_;
};

// MPSstate.h:19-102 (the host-side MPS; its FP64 methods are used for set-up)
class MPSstate {
  public:
    MPSstate(int length, int Dtrunc, int phys_d, Random *RN);
    MPSstate(const char *filename, Random *RN);
    explicit MPSstate(const MPSstate *toCopy);
    MPSstate(int length, int phys_d, const std::vector<int> &vd, const std::vector<double> &data);
    int gLength() const { return L_; }
    int gPhys_d() const { return d_; }
    int gDtrunc() const { return Dtrunc_; }
    int gDimAtBound(int bound) const { return (bound < 0 || bound > L_) ? 0 : vd_[bound]; }
    double LeftNormalize();
    double RightNormalize();
    double InnerProduct(const MPSstate *other) const;
    void ScalarMultiplication(double factor);
    void ChangePhase();
    void CompressState(int truncD);
    void CompressState() { CompressState(Dtrunc_); }
    // this <- H Psi0 (MPSstate.cpp:413-471); bond dims become MPO channels x Psi0's
    void ApplyMPO(const HeisenbergMPO *theMPO, const MPSstate *Psi0);
    // sqrt(s_a) U_a on firstSite and sqrt(s_b) V_b on secondSite, transposed if
    // doHC (MPSstate.cpp:505-573)
    void ApplyTwoSiteTrotterTerm(const TrotterHeisenberg *theTrotter, int firstSite, int secondSite,
                                 int leftSVDindex, int rightSVDindex, bool doHC = false);
    // the grid-point operator sum_k sqrt(s_k dim) x_k U_k (x) V_k (MPSstate.cpp:575-633)
    void ApplyTwoSiteTrotterTerm(const TrotterHeisenberg *theTrotter, int firstSite, int secondSite,
                                 const GridGenerator *theGrid, int gridPoint);
    // exp(dtau h S^z / 2) on every site if the field is on (MPSstate.cpp:635-659)
    void ApplyOneSiteTrotterTermEverywhere(const TrotterHeisenberg *theTrotter);
    const std::vector<int> &bonds() const { return vd_; }
    void assign(const std::vector<int> &vd, const std::vector<double> &data);
    const std::vector<double> &data() const { return data_; }
    std::vector<double> &data() { return data_; }
    friend std::ostream &operator<<(std::ostream &os, const MPSstate &m);

  private:
    long offset(int site) const;
    void apply_site(int site, const double *op);  // A[s'] <- sum_s op[s' + d s] A[s]
    int L_, Dtrunc_, d_;
    std::vector<int> vd_;
    std::vector<double> data_;
};

// DMRG.h:11-60: two-site DMRG for the trial state; Solve() replaces the
// state's tensors by the optimised, left-normalised ones (bond profile of its
// Dtrunc) and returns the energy.  The random start is drawn from a seeded
// stream instead of the state's own random tensors.
class DMRG {
  public:
    DMRG(MPSstate *MPStrial, const HeisenbergMPO *theMPO, unsigned int seed = 20131015u)
        : mps_(MPStrial), mpo_(theMPO), seed_(seed) {}
    double Solve(int maxSweeps = 30, double tol = 1e-10);

  private:
    MPSstate *mps_;
    const HeisenbergMPO *mpo_;
    unsigned int seed_;
};

// Walker.h:13-52: one walker on the host (a copy of a GPU walker, or a walker
// the caller builds and propagates itself with the MPSstate methods)
class Walker {
  public:
    Walker(const MPSstate *state, double overlap, double weight, double energyHistory)
        : state_(state), overlap_(overlap), weight_(weight), eh_(energyHistory) {}
    explicit Walker(const Walker *other)
        : state_(&other->state_), overlap_(other->overlap_), weight_(other->weight_), eh_(other->eh_) {}
    // random MPSstate, weight 1, overlap NAN, energy history 0 (Walker.cpp:28-35)
    Walker(int length, int Dtrunc, int phys_d, Random *RN);
    double gWeight() const { return weight_; }
    double gOverlap() const { return overlap_; }
    double gEnergyHistory() const { return eh_; }
    MPSstate *gState() { return &state_; }
    const MPSstate *gState() const { return &state_; }
    void setWeight(double weight) { weight_ = weight; }
    void multiplyWeight(double factor) { weight_ *= factor; }
    // left-normalise, overlap with the trial, phase fixed so that it is > 0 (Walker.cpp:54-66)
    void setOverlap(const MPSstate *theTrial);
    void setEnergyHistory(double energyHistory) { eh_ = energyHistory; }
    void addEnergyToHistory(double term) { eh_ += term; }
    // <H Psi_T | walker> / overlap; setOverlap must have been called (Walker.cpp:72-77)
    double calcIndividualProjectedEnergy(const MPSstate *HamTimesTrial) const;

  private:
    MPSstate state_;
    double overlap_, weight_, eh_;
};

struct StepRecord {
    int step;
    int nwalkers;
    double projectedEnergy, targetEnergy, fluctMetric, sumWeights;
};

// MPSQMC2.h:18,21,27
class MPSQMC2 {
  public:
    enum class Rng { Host, Device };
    // trial by DMRG with bond dimension DT (MPSQMC2.cpp:225-236)
    MPSQMC2(HeisenbergMPO *theMPO, GridGenerator *theGrid, Random *RN, int DT, int DW, int Nwalkers, double dtau,
            int precision = MQE_PREC_FP64, Rng rng = Rng::Host);
    // trial given
    MPSQMC2(HeisenbergMPO *theMPO, GridGenerator *theGrid, Random *RN, MPSstate *Psi0_in, int DW, int Nwalkers,
            double dtau, int precision = MQE_PREC_FP64, Rng rng = Rng::Host);
    ~MPSQMC2();
    MPSQMC2(const MPSQMC2 &) = delete;
    MPSQMC2 &operator=(const MPSQMC2 &) = delete;

    // Let the walkers propagate for steps steps (MPSQMC2.cpp:443-513)
    void Walk(int steps);

    // Walker population snapshot and exact resume (not in the reference): the
    // next Walk continues from the saved step.  With the host random stream the
    // caller restores that stream itself.
    void saveCheckpoint(const std::string &path) const;
    void loadCheckpoint(const std::string &path);

    // trace file: the reference writes output/Heisenberg1D/ener_L%dDT%dDW%d.txt (MPSQMC2.cpp:452,752)
    void setTraceFile(const std::string &path) { trace_path_ = path; }
    void setVerbose(bool v) { verbose_ = v; }
    const std::vector<StepRecord> &history() const { return history_; }
    double energyAtStart() const { return E_start_; }
    double trialEnergy() const { return E_dmrg_; }  // DMRG energy (DT constructor), else 0
    int nWalkers() const;
    Walker walker(int i) const;
    mqe_engine *engine() { return eng_; }

  private:
    void setup(MPSstate *Psi0_in, int precision);
    double PropagateSeparately();
    void SeparatePopulationControl(double scaling);
    double EnergyFunctionAndHistory(int step, double *projectedEnergy, bool doFluctuationMetric = true);
    void write(int step, int nwalkers, double projectedEnergy, double targetEnergy, double fluctMetric);

    HeisenbergMPO *theMPO_;
    GridGenerator *theGrid_;
    Random *RN_;
    TrotterHeisenberg *theTrotter_;
    MPSstate *Psi0_;
    int DT_, DW_, totalNDesiredWalkers_, totalNCurrentWalkers_, nCouplings_;
    double dtau_;
    Rng rng_;
    mqe_engine *eng_ = nullptr;
    std::string trace_path_;
    bool verbose_ = false;
    bool started_ = false;
    int step_offset_ = 0;
    double E_start_ = 0.0;
    double E_dmrg_ = 0.0;
    std::vector<StepRecord> history_;
};

}  // namespace mpsqmc_amd

#endif
