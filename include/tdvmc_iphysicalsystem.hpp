/* tdvmc_iphysicalsystem.hpp -- the GPU library as a plug-in of the
 * reference driver: PhysicalSystems::GpuBulkBosons derives from the
 * reference's PhysicalSystems::IPhysicalSystem
 * (src/PhysicalSystems/IPhysicalSystem.h:16-123) and implements every one of
 * its virtual functions (:98-122) by forwarding to the C ABI (tdvmc.h) with
 * one configuration at a time (W = 1).
 *
 * Compile it inside the reference's source tree (its include path must reach
 * src/, where PhysicalSystems/IPhysicalSystem.h and the Observables headers
 * lie) and link with -ltdvmc.  It takes the place of BosonsBulk in the
 * driver's factory (src/TDVMC.cpp:2786-2998):
 *
 *     sys = new PhysicalSystems::GpuBulkBosons(SYSTEM_PARAMS, configDirectory, configFilePath);
 *
 * The per-walker interface is there so the driver runs unchanged; the batched
 * entry points (tdvmc_sample, tdvmc_tdvp_step) are the fast path.
 *
 * Semantics against BosonsBulk (src/PhysicalSystems/BosonsBulk.cpp):
 *  - CalculateWavefunction: exponent (with the outer-pair term of the last
 *    pair parameter) and wf = exp(exponent + phiR); O_k kept as localOperators;
 *  - CalculateWFChange / CalculateWFQuotient: R holds the moved particle's new
 *    position; exponentNew, wfNew; quotient exp(2 (exponentNew - exponent))
 *    against the exponent of the last CalculateWavefunction / AcceptMove;
 *  - AcceptMove: the new values become current;
 *  - CalculateExpectationValues: E_R, E_I, O, the products O O^T, O E_R,
 *    O E_I, and the nine 'other' values in the order of :449-457;
 *  - CalculateOtherLocalOperators: the per-particle gradient and Laplacian of
 *    ln psi (the quantities BosonsBulk builds sD / sD2 for), kept for
 *    GetGradient / GetLaplacian;
 *  - CalculateAdditionalSystemProperties: g(r) and S(k) of one configuration
 *    in the two observables BosonsBulk registers (:142-155);
 *  - correlated-sampling data: the kept configuration R, evaluated again at
 *    the current parameters (the reference keeps sD / sD2 instead; with its
 *    reweighting commented out, src/TDVMC.cpp:1245-1256, both give the same
 *    expectation values);
 *  - SetTime (non-virtual in the base) is picked up before every evaluation,
 *    so the quenched pair potential follows the time as in :236-243.
 */
#pragma once

#include "PhysicalSystems/IPhysicalSystem.h"
#include "Observables/ObservableVsOnGrid.h"
#include "Observables/ObservableVsOnGridWithScaling.h"

#include "tdvmc.hpp"

#include <cmath>
#include <memory>
#include <string>
#include <vector>

namespace PhysicalSystems {

// A kept sample: the configuration only (the base class's R, wf, exponent and
// localOperators, which the driver fills itself, src/TDVMC.cpp:1082-1089).
class CSDataGpu : public ICorrelatedSamplingData {
};

class GpuBulkBosons : public IPhysicalSystem {
public:
    // config: path or text of the config file the driver read (its own keys
    // plus this library's); params: SYSTEM_PARAMS as the factory passes them
    GpuBulkBosons(vector<double>& params, string configDirectory, const string& config)
        : IPhysicalSystem(params, configDirectory), ctx_(new tdvmc::Context(config))
    {
        const tdvmc_info& in = ctx_->info();
        this->USE_NORMALIZATION_AND_PHASE = true;
        this->USE_NIC = in.L > 0;
        this->USE_MOVE_COM_TO_ZERO = false;
    }

    tdvmc::Context& context() { return *ctx_; }
    // per-particle grad ln psi [N][DIM] and Laplacian [N] (real parts) of the
    // last CalculateOtherLocalOperators
    const vector<double>& GetGradient() const { return grad_; }
    const vector<double>& GetLaplacian() const { return lap_; }

    // src/PhysicalSystems/BosonsBulk.cpp:51-156: counts of the expectation
    // values and the g(r) / S(k) observables
    void InitSystem() override
    {
        numOfOtherExpectationValues = 9;
        numOfAdditionalSystemProperties = 9;
        const tdvmc_info& in = ctx_->info();
        if (in.L > 0 && in.gr_bin_count > 0) {
            vector<double> grg, skg;
            ctx_->observables_grid(grg, skg);
            const int Ps = in.P_pair;
            const double gmax = ((double)(Ps - 1) * in.rc) / ((double)Ps - 1);
            auto* g = new Observables::ObservableVsOnGridWithScaling();
            g->name = "pairDistribution";
            g->grid.name = "r_ij";
            g->InitGrid(0.0, gmax, gmax / in.gr_bin_count);
            g->InitScaling();
            g->InitObservables({"g_2(r_ij)"});
            auto* s = new Observables::ObservableVsOnGrid();
            s->name = "structureFactor";
            s->grid.name = "k";
            s->InitGrid(skg);
            s->InitObservables({"S(k)"});
            // heap objects: the base destructor deletes what the collection holds
            additionalObservables.Add(g);
            additionalObservables.Add(s);
            gr_ = g;
            sk_ = s;
        }
    }

    // src/PhysicalSystems/BosonsBulk.cpp:338-342
    vector<double> GetCenterOfMass(vector<vector<double> >& R) override
    {
        (void)R;
        return {0.0, 0.0, 0.0};
    }

    void CalculateOtherLocalOperators(vector<vector<double> >& R) override
    {
        flatten(R);
        const int N = ctx_->N(), D = ctx_->DIM(), P = ctx_->P();
        vector<double> ex(1), er(1), ei(1), o(P), g((size_t)N * D * 2), l((size_t)N * 2);
        check(tdvmc_eval_configs(ctx_->handle(), flat_.data(), 1, TDVMC_FP64, ex.data(), er.data(), ei.data(),
                                 o.data(), g.data(), l.data(), nullptr));
        grad_.resize((size_t)N * D);
        lap_.resize(N);
        for (int i = 0; i < N * D; i++) grad_[i] = g[2 * (size_t)i];
        for (int i = 0; i < N; i++) lap_[i] = l[2 * (size_t)i];
    }

    void CalculateExpectationValues(vector<vector<double> >& R, vector<double>& uR, vector<double>& uI, double phiR,
                                    double phiI) override
    {
        eval(R, uR, uI, phiR, phiI);
        exponent = ex_;
        wf = std::exp(exponent + phiR);
        localOperators = O_;
        localEnergyR = ER_;
        localEnergyI = EI_;
        const size_t P = O_.size();
        localOperatorsMatrix.assign(P, vector<double>(P));
        localOperatorlocalEnergyR.resize(P);
        localOperatorlocalEnergyI.resize(P);
        for (size_t i = 0; i < P; i++) {
            for (size_t j = 0; j < P; j++) localOperatorsMatrix[i][j] = O_[i] * O_[j];
            localOperatorlocalEnergyR[i] = O_[i] * ER_;
            localOperatorlocalEnergyI[i] = O_[i] * EI_;
        }
        otherExpectationValues = other_;
    }

    void CalculateExpectationValues(ICorrelatedSamplingData* sample, vector<double>& uR, vector<double>& uI,
                                    double phiR, double phiI) override
    {
        CalculateExpectationValues(sample->R, uR, uI, phiR, phiI);
    }

    // src/PhysicalSystems/BosonsBulk.cpp:474-520 for one configuration
    void CalculateAdditionalSystemProperties(vector<vector<double> >& R, vector<double>& uR, vector<double>& uI,
                                             double phiR, double phiI) override
    {
        (void)uR, (void)uI, (void)phiR, (void)phiI;
        additionalObservables.ClearValues();
        if (!gr_ || !sk_) return;
        flatten(R);
        const size_t ng = gr_->observablesV[0].values.size(), ns = sk_->observablesV[0].values.size();
        vector<double> gr(ng), sk(ns);
        check(tdvmc_observables_configs(ctx_->handle(), flat_.data(), 1, TDVMC_FP64, gr.data(), sk.data()));
        for (size_t b = 0; b < ng; b++) gr_->SetValueAtGridIndex(0, (double)b, gr[b]);
        for (size_t k = 0; k < ns; k++) sk_->SetValueAtGridIndex(0, (double)k, sk[k]);
    }

    // src/PhysicalSystems/BosonsBulk.cpp:522-545
    void CalculateWavefunction(vector<vector<double> >& R, vector<double>& uR, vector<double>& uI, double phiR,
                               double phiI) override
    {
        eval(R, uR, uI, phiR, phiI);
        exponent = ex_;
        wf = std::exp(exponent + phiR);
        localOperators = O_;
    }

    void CalculateWavefunction(ICorrelatedSamplingData* sample, vector<double>& uR, vector<double>& uI, double phiR,
                               double phiI) override
    {
        CalculateWavefunction(sample->R, uR, uI, phiR, phiI);
    }

    // src/PhysicalSystems/BosonsBulk.cpp:553-643: R holds the new position of
    // the changed particle; the whole configuration is evaluated (exactly the
    // incremental update's value, without its running sums)
    void CalculateWFChange(vector<vector<double> >& R, vector<double>& uR, vector<double>& uI, double phiR,
                           double phiI, int changedParticleIndex, vector<double>& oldPosition) override
    {
        (void)changedParticleIndex, (void)oldPosition;
        eval(R, uR, uI, phiR, phiI);
        exponentNew = ex_;
        wfNew = std::exp(exponentNew + phiR);
        Onew_ = O_;
    }

    // src/PhysicalSystems/BosonsBulk.cpp:645-657
    double CalculateWFQuotient(vector<vector<double> >& R, vector<double>& uR, vector<double>& uI, double phiR,
                               double phiI, int changedParticleIndex, vector<double>& oldPosition) override
    {
        CalculateWFChange(R, uR, uI, phiR, phiI, changedParticleIndex, oldPosition);
        return std::exp(2.0 * (exponentNew - exponent));
    }

    // src/PhysicalSystems/BosonsBulk.cpp:659-665
    void AcceptMove() override
    {
        wf = wfNew;
        exponent = exponentNew;
        localOperators = Onew_;
    }

    void InitCorrelatedSamplingData(vector<ICorrelatedSamplingData*>& data) override
    {
        for (auto& d : data) d = new CSDataGpu();
    }

    void FillCorrelatedSamplingData(ICorrelatedSamplingData* data) override { (void)data; }

private:
    void check(int rc)
    {
        if (rc != TDVMC_OK) throw tdvmc::Error(rc, tdvmc_last_error(ctx_->handle()));
    }
    void flatten(const vector<vector<double> >& R)
    {
        flat_.clear();
        for (const auto& r : R) flat_.insert(flat_.end(), r.begin(), r.end());
    }
    // one configuration at the given parameters and the current time
    void eval(vector<vector<double> >& R, vector<double>& uR, vector<double>& uI, double phiR, double phiI)
    {
        if (uR != lastR_ || uI != lastI_ || phiR != lastPhiR_ || phiI != lastPhiI_) {
            ctx_->set_params(uR, uI, phiR, phiI);
            lastR_ = uR;
            lastI_ = uI;
            lastPhiR_ = phiR;
            lastPhiI_ = phiI;
        }
        if (time != lastTime_) {
            ctx_->set_time(time);
            lastTime_ = time;
        }
        flatten(R);
        vector<double> ex, er, ei;
        ctx_->eval_configs(flat_, TDVMC_FP64, ex, er, ei, O_, &other_);
        ex_ = ex[0];
        ER_ = er[0];
        EI_ = ei[0];
    }

    std::unique_ptr<tdvmc::Context> ctx_;
    Observables::ObservableVsOnGridWithScaling* gr_ = nullptr;
    Observables::ObservableVsOnGrid* sk_ = nullptr;
    vector<double> flat_, O_, Onew_, other_, grad_, lap_;
    vector<double> lastR_, lastI_;
    double lastPhiR_ = NAN, lastPhiI_ = NAN, lastTime_ = NAN;
    double ex_ = 0.0, ER_ = 0.0, EI_ = 0.0;
};

}  // namespace PhysicalSystems
