// mpsqmc.cpp -- host classes with the reference's driver interface over the
// engine's C ABI (see include/mpsqmc.hpp).
#include "mpsqmc.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <stdexcept>

namespace mpsqmc_amd {

static void ck(int code, const char *what) {
    if (code != MQE_OK) {
        std::ostringstream os;
        os << what << ": " << mqe_last_error() << " (code " << code << ")";
        throw std::runtime_error(os.str());
    }
}

// ---------------------------------------------------------------- Random
Random::Random(unsigned int seed) : rng_(nullptr), seed_(seed) { ck(mqe_rng_create(seed, &rng_), "Random"); }
Random::~Random() { mqe_rng_destroy(rng_); }
double Random::rand() {
    double u;
    ck(mqe_rng_uniforms(rng_, 1, &u), "Random::rand");
    return u;
}
void Random::fill(long n, double *out) { ck(mqe_rng_uniforms(rng_, n, out), "Random::fill"); }
void Random::peek(long n, double *out) { ck(mqe_rng_peek(rng_, n, out), "Random::peek"); }
void Random::discard(long k) { ck(mqe_rng_discard(rng_, k), "Random::discard"); }

// ---------------------------------------------------------------- GridGenerator
GridGenerator::GridGenerator(int dim) : dim_(dim) {}
void GridGenerator::FillMarsaglia(int pointsPerCo) {
    long cap = dim_;
    for (int c = 0; c < dim_; ++c) cap *= pointsPerCo;
    grid_.assign(cap, 0.0);
    ck(mqe_marsaglia_grid(dim_, pointsPerCo, grid_.data(), &n_), "GridGenerator::FillMarsaglia");
    grid_.resize((size_t)dim_ * n_);
}

// ---------------------------------------------------------------- HeisenbergMPO
HeisenbergMPO::HeisenbergMPO(int length, int phys_d, bool useLadder) : L_(length), d_(phys_d) {
    if (useLadder) throw std::invalid_argument("HeisenbergMPO: the ladder-operator form is not provided");
    J_.assign((size_t)L_ * (L_ - 1) / 2, 0.0);
}
void HeisenbergMPO::sCoupling(int i, int j, double value) {
    if (i == j || i < 0 || j < 0 || i >= L_ || j >= L_) throw std::out_of_range("HeisenbergMPO::sCoupling");
    if (i > j) std::swap(i, j);
    J_[(size_t)(j * (j - 1)) / 2 + i] = value;
}
double HeisenbergMPO::gCoupling(int i, int j) const {
    if (i == j || i < 0 || j < 0 || i >= L_ || j >= L_) return NAN;  // HeisenbergMPO.cpp:278-281
    if (i > j) std::swap(i, j);
    return J_[(size_t)(j * (j - 1)) / 2 + i];
}

// ---------------------------------------------------------------- TrotterHeisenberg
TrotterHeisenberg::TrotterHeisenberg(const HeisenbergMPO *mpo, double dtau)
    : mpo_(mpo), dtau_(dtau), field_(mpo->gField()) {
    const int L = mpo->gLength();
    for (int i = 0; i < L; ++i)
        for (int j = i + 1; j < L; ++j) {
            const double v = mpo->gCoupling(i, j);
            if (v != 0.0 && std::find(dJ_.begin(), dJ_.end(), v) == dJ_.end()) dJ_.push_back(v);
        }
    const int d = mpo->gPhys_d(), n2 = d * d;
    S_.resize(dJ_.size() * n2);
    U_.resize(dJ_.size() * n2 * n2);
    VT_.resize(dJ_.size() * n2 * n2);
    for (size_t q = 0; q < dJ_.size(); ++q)
        ck(mqe_trotter_tables(d, dJ_[q], dtau, &S_[q * n2], &U_[q * n2 * n2], &VT_[q * n2 * n2]),
           "TrotterHeisenberg");
}
int TrotterHeisenberg::find(double J) const {
    for (size_t q = 0; q < dJ_.size(); ++q)
        if (dJ_[q] == J) return (int)q;
    return -1;
}
double TrotterHeisenberg::gTwoSitePropSVD_Sing(double J, int k) const {
    const int q = find(J), n2 = mpo_->gPhys_d() * mpo_->gPhys_d();
    if (q < 0 || k < 0 || k >= n2) return NAN;  // TrotterHeisenberg.cpp:347-361
    return S_[(size_t)q * n2 + k];
}
double TrotterHeisenberg::gTwoSitePropSVD_Left(double J, int k, int i, int j) const {
    const int q = find(J), d = mpo_->gPhys_d(), n2 = d * d;
    if (q < 0 || k < 0 || k >= n2 || i < 0 || j < 0 || i >= d || j >= d) return NAN;
    return U_[(size_t)q * n2 * n2 + i + d * (j + d * k)];
}
double TrotterHeisenberg::gTwoSitePropSVD_Right(double J, int k, int i, int j) const {
    const int q = find(J), d = mpo_->gPhys_d(), n2 = d * d;
    if (q < 0 || k < 0 || k >= n2 || i < 0 || j < 0 || i >= d || j >= d) return NAN;
    return VT_[(size_t)q * n2 * n2 + k + n2 * (i + d * j)];
}
double TrotterHeisenberg::gSingleSiteProp(int i, int j) const {
    const int d = mpo_->gPhys_d();
    if (i < 0 || j < 0 || i >= d || j >= d) return NAN;
    if (i != j) return 0.0;
    return std::exp(0.5 * field_ * dtau_ * (0.5 * (1 - d) + i));  // TrotterHeisenberg.cpp:86
}
void TrotterHeisenberg::tables(const std::vector<double> &Jc, std::vector<double> &S, std::vector<double> &U,
                               std::vector<double> &VT) const {
    const int d = mpo_->gPhys_d(), n2 = d * d;
    S.clear();
    U.clear();
    VT.clear();
    for (double J : Jc) {
        const int q = find(J);
        if (q < 0) throw std::runtime_error("TrotterHeisenberg::tables: unknown coupling");
        S.insert(S.end(), S_.begin() + (size_t)q * n2, S_.begin() + (size_t)(q + 1) * n2);
        U.insert(U.end(), U_.begin() + (size_t)q * n2 * n2, U_.begin() + (size_t)(q + 1) * n2 * n2);
        VT.insert(VT.end(), VT_.begin() + (size_t)q * n2 * n2, VT_.begin() + (size_t)(q + 1) * n2 * n2);
    }
}

// ---------------------------------------------------------------- MPSstate
MPSstate::MPSstate(int length, int Dtrunc, int phys_d, Random *RN) : L_(length), Dtrunc_(Dtrunc), d_(phys_d) {
    vd_.resize(L_ + 1);
    mqe_bond_dims(L_, Dtrunc_, d_, vd_.data());
    data_.resize(offset(L_));
    // MPStensor::random (MPStensor.cpp:75-79): +-U(0,1)
    for (double &x : data_) {
        const double sgn = (RN && RN->rand() < 0.5) ? -1.0 : 1.0;
        x = sgn * (RN ? RN->rand() : 0.0);
    }
}
MPSstate::MPSstate(const char *filename, Random *) {
    std::ifstream in(filename);
    if (!in) throw std::runtime_error(std::string("MPSstate: cannot open ") + filename);
    in >> L_ >> Dtrunc_ >> d_;
    vd_.resize(L_ + 1);
    for (int i = 0; i <= L_; ++i) {
        int idx;
        in >> idx >> vd_[i];
    }
    data_.resize(offset(L_));
    for (int i = 0; i < L_; ++i) {
        int idx, n;
        in >> idx >> n;
        if (n != vd_[i] * vd_[i + 1] * d_) throw std::runtime_error("MPSstate: inconsistent site size in file");
        for (int j = 0; j < n; ++j) {
            int a, b;
            in >> a >> b >> data_[offset(i) + j];
        }
    }
    if (!in) throw std::runtime_error(std::string("MPSstate: truncated file ") + filename);
}
MPSstate::MPSstate(const MPSstate *t) : L_(t->L_), Dtrunc_(t->Dtrunc_), d_(t->d_), vd_(t->vd_), data_(t->data_) {}
MPSstate::MPSstate(int length, int phys_d, const std::vector<int> &vd, const std::vector<double> &data)
    : L_(length), Dtrunc_(*std::max_element(vd.begin(), vd.end())), d_(phys_d), vd_(vd), data_(data) {}

long MPSstate::offset(int site) const {
    long o = 0;
    for (int k = 0; k < site; ++k) o += (long)vd_[k] * vd_[k + 1] * d_;
    return o;
}

double MPSstate::LeftNormalize() {
    double nrm = 0.0;
    ck(mqe_left_normalize(L_, d_, vd_.data(), data_.data(), &nrm), "MPSstate::LeftNormalize");
    return nrm;
}

double MPSstate::InnerProduct(const MPSstate *o) const {
    // MPSstate.cpp:248-284: E = 1; E <- sum_s B[s]^T (E A[s]), this = A, other = B
    if (o->d_ != d_ || o->L_ != L_) return 0.0;
    std::vector<double> E(1, 1.0), W, E2;
    for (int k = 0; k < L_; ++k) {
        const int la = vd_[k], ra = vd_[k + 1], lb = o->vd_[k], rb = o->vd_[k + 1];
        E2.assign((size_t)rb * ra, 0.0);
        W.assign((size_t)lb * ra, 0.0);
        for (int s = 0; s < d_; ++s) {
            const double *A = data_.data() + offset(k) + (size_t)s * la * ra;
            const double *B = o->data_.data() + o->offset(k) + (size_t)s * lb * rb;
            std::fill(W.begin(), W.end(), 0.0);
            for (int c = 0; c < ra; ++c)
                for (int q = 0; q < la; ++q) {
                    const double a = A[q + (size_t)la * c];
                    if (a == 0.0) continue;
                    for (int r = 0; r < lb; ++r) W[r + (size_t)lb * c] += E[r + (size_t)lb * q] * a;
                }
            for (int c = 0; c < ra; ++c)
                for (int r = 0; r < rb; ++r) {
                    double acc = 0.0;
                    for (int q = 0; q < lb; ++q) acc += B[q + (size_t)lb * r] * W[q + (size_t)lb * c];
                    E2[r + (size_t)rb * c] += acc;
                }
        }
        E.swap(E2);
    }
    return E[0];
}

double MPSstate::RightNormalize() {
    double nrm = 0.0;
    ck(mqe_right_normalize(L_, d_, vd_.data(), data_.data(), &nrm), "MPSstate::RightNormalize");
    return nrm;
}

// MPSstate.cpp:286-305: |factor|^(1/L) on every site, the sign on site 0
void MPSstate::ScalarMultiplication(double factor) {
    const double f = std::pow(std::fabs(factor), 1.0 / L_);
    const double sgn = factor > 0.0 ? 1.0 : (factor < 0.0 ? -1.0 : 0.0);
    for (int k = 0; k < L_; ++k) {
        const double a = (k == 0 ? sgn : 1.0) * f;
        const long o = offset(k), n = (long)vd_[k] * vd_[k + 1] * d_;
        for (long i = 0; i < n; ++i) data_[o + i] *= a;
    }
}

void MPSstate::apply_site(int site, const double *op) {
    const long blk = (long)vd_[site] * vd_[site + 1];
    double *A = data_.data() + offset(site);
    std::vector<double> out((size_t)blk * d_, 0.0);
    for (int up = 0; up < d_; ++up)
        for (int dn = 0; dn < d_; ++dn) {
            const double v = op[up + d_ * dn];
            if (v == 0.0) continue;
            for (long e = 0; e < blk; ++e) out[(size_t)up * blk + e] += v * A[(size_t)dn * blk + e];
        }
    std::copy(out.begin(), out.end(), A);
}

void MPSstate::ApplyTwoSiteTrotterTerm(const TrotterHeisenberg *tr, int firstSite, int secondSite, int a, int b,
                                       bool doHC) {
    const double J = tr->gCoupling(firstSite, secondSite);
    std::vector<double> opL((size_t)d_ * d_), opR((size_t)d_ * d_);
    const double sa = std::sqrt(tr->gTwoSitePropSVD_Sing(J, a)), sb = std::sqrt(tr->gTwoSitePropSVD_Sing(J, b));
    for (int up = 0; up < d_; ++up)
        for (int dn = 0; dn < d_; ++dn) {
            const int i = doHC ? dn : up, j = doHC ? up : dn;
            opL[up + d_ * dn] = sa * tr->gTwoSitePropSVD_Left(J, a, i, j);
            opR[up + d_ * dn] = sb * tr->gTwoSitePropSVD_Right(J, b, i, j);
        }
    apply_site(firstSite, opL.data());
    apply_site(secondSite, opR.data());
}

void MPSstate::ApplyTwoSiteTrotterTerm(const TrotterHeisenberg *tr, int firstSite, int secondSite,
                                       const GridGenerator *grid, int gridPoint) {
    const double J = tr->gCoupling(firstSite, secondSite);
    const int dim = grid->gDim();
    std::vector<double> opL((size_t)d_ * d_, 0.0), opR((size_t)d_ * d_, 0.0);
    for (int k = 0; k < dim; ++k) {
        const double pf = std::sqrt(tr->gTwoSitePropSVD_Sing(J, k) * dim) * grid->gCoOfPoint(gridPoint, k);
        if (pf == 0.0) continue;
        for (int up = 0; up < d_; ++up)
            for (int dn = 0; dn < d_; ++dn) {
                opL[up + d_ * dn] += pf * tr->gTwoSitePropSVD_Left(J, k, up, dn);
                opR[up + d_ * dn] += pf * tr->gTwoSitePropSVD_Right(J, k, up, dn);
            }
    }
    apply_site(firstSite, opL.data());
    apply_site(secondSite, opR.data());
}

void MPSstate::ApplyOneSiteTrotterTermEverywhere(const TrotterHeisenberg *tr) {
    if (!tr->gIsMagneticField()) return;
    std::vector<double> op((size_t)d_ * d_);
    for (int up = 0; up < d_; ++up)
        for (int dn = 0; dn < d_; ++dn) op[up + d_ * dn] = tr->gSingleSiteProp(up, dn);
    for (int k = 0; k < L_; ++k) apply_site(k, op.data());
}

void MPSstate::ApplyMPO(const HeisenbergMPO *mpo, const MPSstate *Psi0) {
    const int L = Psi0->L_;
    std::vector<int> first, second;
    std::vector<double> Jc;
    for (int i = 0; i < L; ++i)
        for (int j = i + 1; j < L; ++j)
            if (mpo->gCoupling(i, j) != 0.0) {
                first.push_back(i);
                second.push_back(j);
                Jc.push_back(mpo->gCoupling(i, j));
            }
    std::vector<int> vo(L + 1);
    long n = 0;
    ck(mqe_apply_mpo(L, Psi0->d_, (int)first.size(), first.data(), second.data(), Jc.data(), mpo->gField(),
                     Psi0->vd_.data(), Psi0->data_.data(), vo.data(), nullptr, &n),
       "MPSstate::ApplyMPO");
    std::vector<double> out(n);
    ck(mqe_apply_mpo(L, Psi0->d_, (int)first.size(), first.data(), second.data(), Jc.data(), mpo->gField(),
                     Psi0->vd_.data(), Psi0->data_.data(), vo.data(), out.data(), &n),
       "MPSstate::ApplyMPO");
    L_ = L;
    d_ = Psi0->d_;
    vd_ = vo;
    data_.swap(out);
    Dtrunc_ = *std::max_element(vd_.begin(), vd_.end());
}

void MPSstate::ChangePhase() {  // MPSstate.cpp:307-315
    const long n = (long)vd_[0] * vd_[1] * d_;
    for (long i = 0; i < n; ++i) data_[i] = -data_[i];
}

void MPSstate::CompressState(int truncD) {
    std::vector<int> vo(L_ + 1);
    std::vector<double> out(data_.size());
    ck(mqe_compress(L_, d_, vd_.data(), data_.data(), truncD, vo.data(), out.data()), "MPSstate::CompressState");
    vd_ = vo;
    out.resize(offset(L_));
    data_.swap(out);
    Dtrunc_ = *std::max_element(vd_.begin(), vd_.end());
}

std::ostream &operator<<(std::ostream &os, const MPSstate &m) {  // MPSstate.cpp:661-686
    os << m.L_ << "\t" << m.Dtrunc_ << "\t" << m.d_ << "\n";
    for (int i = 0; i <= m.L_; ++i) os << i << "\t" << m.vd_[i] << "\n";
    os << std::setprecision(17);
    for (int i = 0; i < m.L_; ++i) {
        const long n = (long)m.vd_[i] * m.vd_[i + 1] * m.d_, o = m.offset(i);
        os << i << "\t" << n << "\n";
        for (long j = 0; j < n; ++j) os << i << "\t" << j << "\t" << m.data_[o + j] << "\n";
    }
    return os;
}

void MPSstate::assign(const std::vector<int> &vd, const std::vector<double> &data) {
    if ((int)vd.size() != L_ + 1) throw std::invalid_argument("MPSstate::assign: bond count");
    vd_ = vd;
    data_.assign(data.begin(), data.begin() + offset(L_));
}

// ---------------------------------------------------------------- Walker
Walker::Walker(int length, int Dtrunc, int phys_d, Random *RN)
    : state_(length, Dtrunc, phys_d, RN), overlap_(NAN), weight_(1.0), eh_(0.0) {}

void Walker::setOverlap(const MPSstate *theTrial) {
    state_.LeftNormalize();
    double v = theTrial->InnerProduct(&state_);
    if (v < 0.0) {  // importance sampling keeps O > 0
        state_.ChangePhase();
        v = -v;
    }
    overlap_ = v;
}

double Walker::calcIndividualProjectedEnergy(const MPSstate *HamTimesTrial) const {
    return HamTimesTrial->InnerProduct(&state_) / overlap_;
}

// ---------------------------------------------------------------- DMRG
double DMRG::Solve(int maxSweeps, double tol) {
    const int L = mpo_->gLength(), d = mpo_->gPhys_d();
    std::vector<int> first, second;
    std::vector<double> Jc;
    for (int i = 0; i < L - 1; ++i)
        for (int j = i + 1; j < L; ++j)
            if (mpo_->gCoupling(i, j) != 0.0) {
                first.push_back(i);
                second.push_back(j);
                Jc.push_back(mpo_->gCoupling(i, j));
            }
    std::vector<int> vd(L + 1);
    mqe_bond_dims(L, mps_->gDtrunc(), d, vd.data());
    long n = 0;
    for (int k = 0; k < L; ++k) n += (long)vd[k] * vd[k + 1] * d;
    std::vector<double> data(n);
    double E = 0.0;
    ck(mqe_dmrg(L, d, (int)first.size(), first.data(), second.data(), Jc.data(), mpo_->gField(), mps_->gDtrunc(),
                maxSweeps, tol, seed_, vd.data(), data.data(), &E),
       "DMRG::Solve");
    mps_->assign(vd, data);
    return E;
}

// ---------------------------------------------------------------- MPSQMC2
MPSQMC2::MPSQMC2(HeisenbergMPO *theMPO, GridGenerator *theGrid, Random *RN, int DT, int DW, int Nwalkers,
                 double dtau, int precision, Rng rng)
    : theMPO_(theMPO), theGrid_(theGrid), RN_(RN), DW_(DW), totalNDesiredWalkers_(Nwalkers),
      totalNCurrentWalkers_(Nwalkers), dtau_(dtau), rng_(rng) {
    // SetupTrial(true) (MPSQMC2.cpp:225-236): DMRG, then left-normalise
    MPSstate trial(theMPO->gLength(), DT, theMPO->gPhys_d(), nullptr);
    DMRG solver(&trial, theMPO);
    E_dmrg_ = solver.Solve();
    setup(&trial, precision);
}

MPSQMC2::MPSQMC2(HeisenbergMPO *theMPO, GridGenerator *theGrid, Random *RN, MPSstate *Psi0_in, int DW,
                 int Nwalkers, double dtau, int precision, Rng rng)
    : theMPO_(theMPO), theGrid_(theGrid), RN_(RN), DW_(DW), totalNDesiredWalkers_(Nwalkers),
      totalNCurrentWalkers_(Nwalkers), dtau_(dtau), rng_(rng) {
    setup(Psi0_in, precision);
}

void MPSQMC2::setup(MPSstate *Psi0_in, int precision) {
    HeisenbergMPO *theMPO = theMPO_;
    GridGenerator *theGrid = theGrid_;
    const int DW = DW_, Nwalkers = totalNDesiredWalkers_;
    const double dtau = dtau_;
    const int L = theMPO->gLength(), d = theMPO->gPhys_d();
    if (theGrid->gDim() != d * d) throw std::invalid_argument("MPSQMC2: grid dimension must be d*d");
    theTrotter_ = new TrotterHeisenberg(theMPO, dtau);
    // SetupTrial (MPSQMC2.cpp:219-332): left-normalised copy of the trial
    Psi0_ = new MPSstate(Psi0_in);
    Psi0_->LeftNormalize();
    DT_ = Psi0_->gDtrunc();
    std::vector<int> first, second;
    std::vector<double> Jc;
    for (int i = 0; i < L - 1; ++i)
        for (int j = i + 1; j < L; ++j)
            if (theMPO->gCoupling(i, j) != 0.0) {
                first.push_back(i);
                second.push_back(j);
                Jc.push_back(theMPO->gCoupling(i, j));
            }
    nCouplings_ = (int)first.size();
    std::vector<double> S, U, VT;
    theTrotter_->tables(Jc, S, U, VT);
    // SetupWalkers (MPSQMC2.cpp:381-441), copyTrial = true
    MPSstate w0(Psi0_);
    double O = 1.0;
    if (DW < DT_) {
        w0.CompressState(DW);
        w0.LeftNormalize();
        O = Psi0_->InnerProduct(&w0);
        if (O < 0.0) {
            w0.ChangePhase();
            O = -O;
        }
    }
    mqe_params prm;
    prm.L = L;
    prm.d = d;
    prm.precision = precision;
    prm.n_desired = Nwalkers;
    prm.capacity = std::max(1000, 3 * Nwalkers);  // MPSQMC2.cpp:84
    prm.max_bond = std::min(DW, DT_);
    prm.dtau = dtau;
    prm.field = theMPO->gField();
    prm.seed = RN_ ? (unsigned long long)RN_->seed() : 20131015ull;  // device draws follow the caller's seed
    mqe_mps t;
    t.L = L;
    t.d = d;
    t.vd = Psi0_->bonds().data();
    t.data = Psi0_->data().data();
    ck(mqe_create(&prm, &t, nCouplings_, first.data(), second.data(), Jc.data(), S.data(), U.data(), VT.data(),
                  theGrid->data(), theGrid->gNPoints(), theGrid->gDim(), &eng_),
       "MPSQMC2: mqe_create");
    std::vector<int> vds;
    std::vector<double> datas;
    for (int i = 0; i < Nwalkers; ++i) {
        vds.insert(vds.end(), w0.bonds().begin(), w0.bonds().end());
        datas.insert(datas.end(), w0.data().begin(), w0.data().end());
    }
    std::vector<double> w(Nwalkers, 1.0), Ov(Nwalkers, O), eh(Nwalkers, 0.0);
    ck(mqe_load_walkers(eng_, Nwalkers, vds.data(), datas.data(), w.data(), Ov.data(), eh.data()),
       "MPSQMC2: mqe_load_walkers");
}

MPSQMC2::~MPSQMC2() {
    mqe_destroy(eng_);
    delete theTrotter_;
    delete Psi0_;
}

double MPSQMC2::PropagateSeparately() {
    double sw = 0.0;
    if (rng_ == Rng::Host) {
        std::vector<double> u((size_t)totalNCurrentWalkers_ * nCouplings_);
        RN_->fill((long)u.size(), u.data());
        ck(mqe_propagate(eng_, u.data(), &sw), "PropagateSeparately");
    } else {
        ck(mqe_propagate(eng_, nullptr, &sw), "PropagateSeparately");
    }
    return sw;
}

void MPSQMC2::SeparatePopulationControl(double scaling) {
    int nn = 0, used = 0;
    if (rng_ == Rng::Host) {
        std::vector<double> u(totalNCurrentWalkers_);
        RN_->peek((long)u.size(), u.data());
        ck(mqe_population_control(eng_, scaling, u.data(), (int)u.size(), &nn, &used), "SeparatePopulationControl");
        RN_->discard(used);
    } else {
        ck(mqe_population_control(eng_, scaling, nullptr, 0, &nn, &used), "SeparatePopulationControl");
    }
    totalNCurrentWalkers_ = nn;
}

double MPSQMC2::EnergyFunctionAndHistory(int step, double *projectedEnergy, bool doFluct) {
    double om = 0.0;
    ck(mqe_energy(eng_, step, doFluct ? 1 : 0, projectedEnergy, &om), "EnergyFunctionAndHistory");
    return om;
}

void MPSQMC2::write(int step, int nwalkers, double EP, double ET, double om) {
    if (trace_path_.empty()) return;
    std::ofstream out(trace_path_, std::ios::app);
    out.precision(10);
    out << step << "\t\t" << nwalkers << "\t" << EP << "\t\t" << ET << "\t\t" << om << std::endl;
}

int MPSQMC2::nWalkers() const { return totalNCurrentWalkers_; }

Walker MPSQMC2::walker(int i) const {
    const int L = theMPO_->gLength();
    std::vector<int> vd(L + 1);
    long n = 0;
    ck(mqe_read_walker(eng_, i, vd.data(), nullptr, &n), "walker");
    std::vector<double> data(n);
    ck(mqe_read_walker(eng_, i, vd.data(), data.data(), &n), "walker");
    std::vector<double> w(totalNCurrentWalkers_), O(totalNCurrentWalkers_), eh(totalNCurrentWalkers_);
    ck(mqe_read_scalars(eng_, w.data(), O.data(), eh.data(), nullptr), "walker");
    MPSstate st(L, theMPO_->gPhys_d(), vd, data);
    return Walker(&st, O[i], w[i], eh[i]);
}

void MPSQMC2::saveCheckpoint(const std::string &path) const {
    ck(mqe_save_state(eng_, path.c_str(), step_offset_), "MPSQMC2::saveCheckpoint");
}

void MPSQMC2::loadCheckpoint(const std::string &path) {
    long st = 0;
    ck(mqe_load_state(eng_, path.c_str(), &st), "MPSQMC2::loadCheckpoint");
    int n = 0;
    ck(mqe_num_walkers(eng_, &n), "MPSQMC2::loadCheckpoint");
    totalNCurrentWalkers_ = n;
    step_offset_ = (int)st;
    started_ = true;
}

void MPSQMC2::Walk(int steps) {
    double projectedEnergy = 0.0;
    if (!started_) {
        ck(mqe_prepare(eng_, &projectedEnergy), "Walk: prepare");  // MPSQMC2.cpp:447
        E_start_ = projectedEnergy;
        if (!trace_path_.empty()) {
            std::ofstream out(trace_path_, std::ios::trunc);
            out << "#Step\t\tN\t\tE_P\t\tE_T\t\tFluctMetric" << std::endl;
        }
        if (verbose_) std::cout << "Energy at start = " << projectedEnergy << std::endl;
        started_ = true;
    }
    for (int s = 1; s <= steps; ++s) {
        const int step = step_offset_ + s;
        const double sumW = PropagateSeparately();
        const double avgWeight = sumW / totalNCurrentWalkers_;
        const double scaling = totalNDesiredWalkers_ / (totalNCurrentWalkers_ * avgWeight);
        const double targetEnergy = std::log(scaling) / dtau_;
        const double fluct = EnergyFunctionAndHistory(step, &projectedEnergy, true);
        if (verbose_)
            std::cout << "step " << step << "  N " << totalNCurrentWalkers_ << "  E_P " << projectedEnergy << "  E_T "
                      << targetEnergy << "  Omega " << fluct << std::endl;
        write(step, totalNCurrentWalkers_, projectedEnergy, targetEnergy, fluct);
        history_.push_back({step, totalNCurrentWalkers_, projectedEnergy, targetEnergy, fluct, sumW});
        SeparatePopulationControl(scaling);
    }
    step_offset_ += steps;
}

}  // namespace mpsqmc_amd
