// The plain Newton optimiser over per-operator angles (see newton.cpp).
#pragma once
#include <chrono>
#include <functional>
#include <vector>

namespace ae {

void symmetric_eigen(int n, std::vector<double> A, std::vector<double> &w, std::vector<double> &V);

struct EvalFns {
    // per-operator angles -> per-parameter Hessian (row-major), gradient, E
    std::function<bool(const std::vector<double> &, std::vector<double> &, std::vector<double> &, double &)> hessian;
    // per-operator angles -> E and per-parameter gradient
    std::function<bool(const std::vector<double> &, double &, std::vector<double> &)> egrad;
};

struct TracePoint {
    double seconds, energy, gradNorm;
    long long hessianEvals, energyEvals;
};

class NewtonDriver {
public:
    NewtonDriver(const std::vector<int> &order, const std::vector<double> &scale, int nparams, EvalFns fns);
    // Runs OptimiseTups to ||g|| < gradTolerance (or until stuck twice, or
    // until the energy reaches stopEnergy when set).  angles: per operator.
    bool optimise(std::vector<double> &angles, double &E);
    void set_stop_energy(double e)
    {
        stopEnergy_ = true;
        stopEnergyValue_ = e;
    }

    double gradTolerance = 1e-12;
    int maxNewtonSteps = 500;
    int maxOuter = 50;
    long long hessianEvals = 0, energyEvals = 0;
    std::vector<TracePoint> trace;

private:
    bool newton(std::vector<double> &angles);
    bool escape(std::vector<double> &angles, bool &foundAny);
    std::vector<double> decompress(const std::vector<double> &step) const;
    void record(double E, double gnorm);

    std::vector<int> order_;
    std::vector<double> scale_;
    int P_, M_;
    EvalFns f_;
    bool stopEnergy_ = false;
    double stopEnergyValue_ = 0;
    std::chrono::steady_clock::time_point t0_;
};

} // namespace ae
