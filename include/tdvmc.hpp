/* tdvmc.hpp -- header-only C++ layer over the C ABI of tdvmc.h.
 *
 * Three pieces, for C++ callers shaped like the reference's driver:
 *   tdvmc::Context            RAII owner of a tdvmc_ctx (one HIP stream, all
 *                             device buffers); errors become tdvmc::Error.
 *   tdvmc::SimulationStepData the fields of the reference's SimulationStepData
 *                             (src/SimulationStepData.h:7-31), filled from
 *                             one batched sampling stage.
 *   tdvmc::PhysicalSystem     the per-walker calls of
 *                             PhysicalSystems::IPhysicalSystem
 *                             (src/PhysicalSystems/IPhysicalSystem.h:16-123)
 *                             for one configuration at a time, evaluated
 *                             through tdvmc_eval_configs with W = 1.  It is
 *                             there for interface compatibility and parity
 *                             checks; the batched calls of Context are the
 *                             fast path.
 *
 * Requires only tdvmc.h and the C++ standard library; link with -ltdvmc.
 */
#pragma once

#include "tdvmc.h"

#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace tdvmc {

class Error : public std::runtime_error {
public:
    Error(int code, const std::string& msg) : std::runtime_error(msg), code_(code) {}
    int code() const { return code_; }

private:
    int code_;
};

// Fields of src/SimulationStepData.h:7-31; matrices as nested vectors like
// the reference, plus the centred moments S and F the solve consumes.
struct SimulationStepData {
    double time = 0.0;
    int step = 0;
    double phiR = 0.0, phiI = 0.0;
    std::vector<double> paramsR, paramsI;
    double phiRDiff = 0.0, phiIDiff = 0.0;
    std::vector<double> paramsRDiff, paramsIDiff;
    std::vector<double> localOperators;
    double localEnergyR = 0.0, localEnergyI = 0.0;
    std::vector<std::vector<double>> localOperatorsMatrix;
    std::vector<double> localOperatorlocalEnergyR, localOperatorlocalEnergyI;
    std::vector<double> otherExpectationValues;
    std::vector<std::vector<double>> S;
    std::vector<double> FR, FI;
    int64_t n_samples = 0;
};

class Context {
public:
    // config: the text of a config/NAME.config JSON document or a path to one
    explicit Context(const std::string& config)
    {
        check(tdvmc_create(config.c_str(), &ctx_), nullptr);
        check(tdvmc_get_info(ctx_, &info_));
    }
    ~Context()
    {
        if (ctx_) tdvmc_destroy(ctx_);
    }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;

    tdvmc_ctx* handle() { return ctx_; }
    const tdvmc_info& info()
    {
        check(tdvmc_get_info(ctx_, &info_));
        return info_;
    }
    int N() const { return info_.N; }
    int DIM() const { return info_.DIM; }
    int P() const { return info_.P; }

    void set_walker_count(int64_t W) { check(tdvmc_set_walker_count(ctx_, W)); }
    void set_params(const std::vector<double>& uR, const std::vector<double>& uI, double phiR, double phiI)
    {
        need(uR.size() == (size_t)P() && uI.size() == (size_t)P(), "parameter vectors must have P entries");
        check(tdvmc_set_params(ctx_, uR.data(), uI.data(), phiR, phiI, P()));
    }
    void get_params(std::vector<double>& uR, std::vector<double>& uI, double& phiR, double& phiI)
    {
        uR.resize(P());
        uI.resize(P());
        check(tdvmc_get_params(ctx_, uR.data(), uI.data(), &phiR, &phiI, P()));
    }
    void set_time(double t) { check(tdvmc_set_time(ctx_, t)); }
    // walker coordinates, [W][N][DIM] row-major
    void set_walkers(const std::vector<double>& R)
    {
        const size_t per = (size_t)N() * DIM();
        need(per > 0 && R.size() % per == 0, "R must hold whole walkers");
        check(tdvmc_set_walkers(ctx_, R.data(), (int64_t)(R.size() / per)));
    }
    std::vector<double> get_walkers()
    {
        std::vector<double> R((size_t)info().n_walkers * N() * DIM());
        check(tdvmc_get_walkers(ctx_, R.data(), info_.n_walkers));
        return R;
    }
    void init_walkers(uint64_t seed) { check(tdvmc_init_walkers(ctx_, seed)); }
    void sweep(int64_t n_moves, uint64_t seed, int precision = TDVMC_PRECISION_DEFAULT)
    {
        check(tdvmc_sweep(ctx_, n_moves, seed, precision));
    }
    // one sampling stage (UpdateExpectationValues, src/TDVMC.cpp:1038-1220)
    SimulationStepData sample(uint64_t seed, int precision = TDVMC_PRECISION_DEFAULT)
    {
        return step_data([&](tdvmc_step_data* d) { return tdvmc_sample(ctx_, seed, precision, d); });
    }
    // measurement of the current walkers only
    SimulationStepData measure(int precision = TDVMC_PRECISION_DEFAULT)
    {
        return step_data([&](tdvmc_step_data* d) { return tdvmc_measure(ctx_, precision, d); });
    }
    // CalculateNextParameters (src/TDVMC.cpp:1834-1853, :1969-2035) on the device
    void tdvp_step(double dt, int ode_type = -1, int precision = TDVMC_PRECISION_DEFAULT, bool graph = true)
    {
        check(tdvmc_tdvp_step(ctx_, dt, ode_type, precision, graph ? 1 : 0));
    }
    // ln psi, E_L, O for fed configurations R [W][N][DIM]; other [W][9]
    void eval_configs(const std::vector<double>& R, int precision, std::vector<double>& expo,
                      std::vector<double>& ER, std::vector<double>& EI, std::vector<double>& O,
                      std::vector<double>* other = nullptr)
    {
        const size_t per = (size_t)N() * DIM();
        need(per > 0 && R.size() % per == 0, "R must hold whole walkers");
        const int64_t W = (int64_t)(R.size() / per);
        expo.resize(W);
        ER.resize(W);
        EI.resize(W);
        O.resize((size_t)W * P());
        if (other) other->resize((size_t)W * 9);
        check(tdvmc_eval_configs(ctx_, R.data(), W, precision, expo.data(), ER.data(), EI.data(), O.data(),
                                 nullptr, nullptr, other ? other->data() : nullptr));
    }
    // grids of the pair distribution (bin centres) and S(k) (|k| per shell)
    void observables_grid(std::vector<double>& gr_grid, std::vector<double>& sk_grid)
    {
        int ng = 0, ns = 0;
        check(tdvmc_observables_info(ctx_, &ng, &ns, nullptr, nullptr));
        gr_grid.resize(ng);
        sk_grid.resize(ns);
        check(tdvmc_observables_info(ctx_, nullptr, nullptr, gr_grid.data(), sk_grid.data()));
    }
    // ParallelCalculateAdditionalSystemProperties (src/TDVMC.cpp:1332-1365, :1438-1444):
    // g(r) and S(k) averaged over n_samples observations (<= 0: MC_NADDITIONALSTEPS) of every walker
    void additional_properties(uint64_t seed, std::vector<double>& gr, std::vector<double>& sk,
                               int64_t n_samples = 0, int precision = TDVMC_PRECISION_DEFAULT)
    {
        int ng = 0, ns = 0;
        check(tdvmc_observables_info(ctx_, &ng, &ns, nullptr, nullptr));
        gr.resize(ng);
        sk.resize(ns);
        check(tdvmc_additional_properties(ctx_, n_samples, seed, precision, gr.data(), sk.data()));
    }

private:
    template <typename F>
    SimulationStepData step_data(F call)
    {
        const int P = this->P();
        SimulationStepData s;
        std::vector<double> M((size_t)P * P), Sf((size_t)P * P);
        s.localOperators.resize(P);
        s.localOperatorlocalEnergyR.resize(P);
        s.localOperatorlocalEnergyI.resize(P);
        s.FR.resize(P);
        s.FI.resize(P);
        s.otherExpectationValues.resize(9);
        tdvmc_step_data d{};
        d.localOperators = s.localOperators.data();
        d.localOperatorsMatrix = M.data();
        d.localOperatorlocalEnergyR = s.localOperatorlocalEnergyR.data();
        d.localOperatorlocalEnergyI = s.localOperatorlocalEnergyI.data();
        d.S = Sf.data();
        d.FR = s.FR.data();
        d.FI = s.FI.data();
        d.otherExpectationValues = s.otherExpectationValues.data();
        check(call(&d));
        s.time = d.time;
        s.step = d.step;
        s.n_samples = d.n_samples;
        s.localEnergyR = d.localEnergyR;
        s.localEnergyI = d.localEnergyI;
        s.localOperatorsMatrix.assign(P, std::vector<double>(P));
        s.S.assign(P, std::vector<double>(P));
        for (int i = 0; i < P; i++)
            for (int j = 0; j < P; j++) {
                s.localOperatorsMatrix[i][j] = M[(size_t)i * P + j];
                s.S[i][j] = Sf[(size_t)i * P + j];
            }
        get_params(s.paramsR, s.paramsI, s.phiR, s.phiI);
        return s;
    }
    void check(int rc, tdvmc_ctx* c)
    {
        if (rc != TDVMC_OK) throw Error(rc, tdvmc_last_error(c));
    }
    void check(int rc) { check(rc, ctx_); }
    void need(bool ok, const char* msg)
    {
        if (!ok) throw Error(TDVMC_ERR_BAD_ARG, msg);
    }

    tdvmc_ctx* ctx_ = nullptr;
    tdvmc_info info_{};
};

// The per-walker surface of IPhysicalSystem (src/PhysicalSystems/IPhysicalSystem.h:76-122)
// for one configuration R[N][DIM] at a time; every evaluation is one
// tdvmc_eval_configs call with W = 1 on the context's device.
class PhysicalSystem {
public:
    explicit PhysicalSystem(Context& ctx) : ctx_(ctx) {}

    double GetTime() { return time_; }
    void SetTime(double t)
    {
        time_ = t;
        ctx_.set_time(t);   // selects the quenched pair potential
    }
    int GetStep() { return step_; }
    void SetStep(double s) { step_ = (int)s; }
    double GetWf() { return wf_; }
    double GetWfNew() { return wfNew_; }
    double GetExponent() { return exponent_; }
    double GetExponentNew() { return exponentNew_; }
    double GetLocalEnergyR() { return localEnergyR_; }
    double GetLocalEnergyI() { return localEnergyI_; }
    std::vector<double> GetLocalOperators() { return localOperators_; }
    std::vector<std::vector<double>> GetLocalOperatorsMatrix() { return localOperatorsMatrix_; }
    std::vector<double> GetLocalOperatorlocalEnergyR() { return localOperatorlocalEnergyR_; }
    std::vector<double> GetLocalOperatorlocalEnergyI() { return localOperatorlocalEnergyI_; }
    std::vector<double> GetOtherExpectationValues() { return otherExpectationValues_; }
    int GetNumOfOtherExpectationValues() { return 9; }

    // src/PhysicalSystems/BosonsBulk.cpp:522-539: exponent and wf = exp(exponent + phiR)
    void CalculateWavefunction(std::vector<std::vector<double>>& R, std::vector<double>& uR,
                               std::vector<double>& uI, double phiR, double phiI)
    {
        eval(R, uR, uI, phiR, phiI);
        exponent_ = expo_[0];
        wf_ = std::exp(exponent_ + phiR);
    }

    // src/PhysicalSystems/BosonsBulk.cpp:349-458: E_L, O and the products the
    // driver averages (O O^T, O E_R, O E_I) for one configuration
    void CalculateExpectationValues(std::vector<std::vector<double>>& R, std::vector<double>& uR,
                                    std::vector<double>& uI, double phiR, double phiI)
    {
        eval(R, uR, uI, phiR, phiI);
        const int P = ctx_.P();
        exponent_ = expo_[0];
        wf_ = std::exp(exponent_ + phiR);
        localEnergyR_ = ER_[0];
        localEnergyI_ = EI_[0];
        localOperators_.assign(O_.begin(), O_.begin() + P);
        localOperatorsMatrix_.assign(P, std::vector<double>(P));
        localOperatorlocalEnergyR_.resize(P);
        localOperatorlocalEnergyI_.resize(P);
        for (int i = 0; i < P; i++) {
            for (int j = 0; j < P; j++) localOperatorsMatrix_[i][j] = O_[i] * O_[j];
            localOperatorlocalEnergyR_[i] = O_[i] * ER_[0];
            localOperatorlocalEnergyI_[i] = O_[i] * EI_[0];
        }
        otherExpectationValues_ = other_;
    }

    // src/PhysicalSystems/BosonsBulk.cpp:553-657: R already holds the moved
    // particle's new position; returns |psi_new / psi_old|^2
    double CalculateWFQuotient(std::vector<std::vector<double>>& R, std::vector<double>& uR,
                               std::vector<double>& uI, double phiR, double phiI, int changedParticleIndex,
                               std::vector<double>& oldPosition)
    {
        std::vector<std::vector<double>> Rold = R;
        Rold[changedParticleIndex] = oldPosition;
        eval(Rold, uR, uI, phiR, phiI);
        exponent_ = expo_[0];
        wf_ = std::exp(exponent_ + phiR);
        eval(R, uR, uI, phiR, phiI);
        exponentNew_ = expo_[0];
        wfNew_ = std::exp(exponentNew_ + phiR);
        return std::exp(2.0 * (exponentNew_ - exponent_));
    }

    // src/PhysicalSystems/BosonsBulk.cpp:659-665
    void AcceptMove()
    {
        exponent_ = exponentNew_;
        wf_ = wfNew_;
    }

private:
    void eval(std::vector<std::vector<double>>& R, std::vector<double>& uR, std::vector<double>& uI, double phiR,
              double phiI)
    {
        if (uR != lastR_ || uI != lastI_ || phiR != lastPhiR_ || phiI != lastPhiI_) {
            ctx_.set_params(uR, uI, phiR, phiI);
            lastR_ = uR;
            lastI_ = uI;
            lastPhiR_ = phiR;
            lastPhiI_ = phiI;
        }
        flat_.clear();
        for (const auto& r : R) flat_.insert(flat_.end(), r.begin(), r.end());
        ctx_.eval_configs(flat_, TDVMC_FP64, expo_, ER_, EI_, O_, &other_);
    }

    Context& ctx_;
    double time_ = 0.0;
    int step_ = 0;
    double wf_ = 0.0, wfNew_ = 0.0, exponent_ = 0.0, exponentNew_ = 0.0;
    double localEnergyR_ = 0.0, localEnergyI_ = 0.0;
    std::vector<double> localOperators_, localOperatorlocalEnergyR_, localOperatorlocalEnergyI_;
    std::vector<std::vector<double>> localOperatorsMatrix_;
    std::vector<double> otherExpectationValues_;
    std::vector<double> lastR_, lastI_, flat_, expo_, ER_, EI_, O_, other_;
    double lastPhiR_ = NAN, lastPhiI_ = NAN;
};

}  // namespace tdvmc
