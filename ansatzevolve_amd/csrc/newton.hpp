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

// Callbacks of the projected optimiser (projected.cpp).  The projection
// vector p is held by the caller.
struct ProjFns {
    // per-operator angles -> psi(angles) over the sector, and p := H psi(angles)
    std::function<bool(const std::vector<double> &, std::vector<double> &)> setProjection;
    // per-operator angles -> per-parameter Hessian (row-major), gradient and
    // value of f = <p|psi(angles)>
    std::function<bool(const std::vector<double> &, std::vector<double> &, std::vector<double> &, double &)>
        hessianProj;
    // per-operator angles -> f and its per-parameter gradient
    std::function<bool(const std::vector<double> &, double &, std::vector<double> &)> gradProj;
};

// TUPSQuantities::iterativeTups(FusedEvolve&) with runNewtonMethodProjected
// (src/tupsquantities.cpp:2044-2145, 1238-1465): outer iterations with the
// projection p = H psi(previous result), each minimising <p|psi(theta)> by
// Newton steps, until the projected gradient's norm is below gradTolerance.
class ProjectedDriver {
public:
    ProjectedDriver(const std::vector<int> &order, const std::vector<double> &scale, int nparams, ProjFns fns);
    // angles: per operator, projected onto the parameter space on entry;
    // E: the energy <H psi|psi> at the result
    bool optimise(std::vector<double> &angles, double &E);

    double gradTolerance = 1e-5;
    int maxNewtonSteps = 500;
    int maxOuter = 3000;
    bool useEdiis = false;  // the reference's default (doDIIS = false)
    int ediisSize = 15;
    long long hessianEvals = 0, energyEvals = 0;
    std::vector<TracePoint> trace;  // one point per outer iteration

private:
    bool newton(std::vector<double> &angles);
    std::vector<double> to_ops(const std::vector<double> &theta) const;
    std::vector<double> to_params(const std::vector<double> &x) const;

    std::vector<int> order_;
    std::vector<double> scale_;
    int P_, M_;
    ProjFns f_;
    std::chrono::steady_clock::time_point t0_;
};

// Energy-DIIS coefficients for the n x n Gram matrix B of the error vectors
// (row-major): c_i >= 0, sum c_i = 1, minimising c^T B c.
std::vector<double> ediis_coefficients(int n, const std::vector<double> &B);

} // namespace ae
