// The plain optimiser: Newton steps with an |lambda|^-1 spectral inverse, a
// cubic-model line search and negative-curvature escape steps, restating
// TUPSQuantities::runNewtonMethod (src/tupsquantities.cpp:704-1236),
// doStepsUntilHessianIsPositiveDefinite (:1467-1598) and
// OptimiseTups(FusedEvolve&) (:1940-2011).  Energies, gradients and
// Hessians come from the device engine; the P x P eigen-decomposition runs
// on the host (P <= a few hundred).
#include "newton.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>

namespace ae {

// Cyclic Jacobi eigen-decomposition of a symmetric n x n matrix (row-major).
// On return w holds the eigenvalues (ascending) and V the eigenvectors as
// columns, V[i*n + k] = component i of eigenvector k.
void symmetric_eigen(int n, std::vector<double> A, std::vector<double> &w, std::vector<double> &V)
{
    V.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++) V[(size_t)i * n + i] = 1.0;
    auto a = [&](int i, int j) -> double & { return A[(size_t)i * n + j]; };
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0, tot = 0;
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                const double v = a(i, j) * a(i, j);
                tot += v;
                if (i != j) off += v;
            }
        if (off <= 1e-30 * tot || off == 0.0) break;
        for (int p = 0; p < n - 1; p++)
            for (int q = p + 1; q < n; q++) {
                const double apq = a(p, q);
                if (std::abs(apq) < 1e-300) continue;
                const double app = a(p, p), aqq = a(q, q);
                const double theta = (aqq - app) / (2.0 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::abs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; k++) {  // columns p, q
                    const double akp = a(k, p), akq = a(k, q);
                    a(k, p) = c * akp - s * akq;
                    a(k, q) = s * akp + c * akq;
                }
                for (int k = 0; k < n; k++) {  // rows p, q
                    const double apk = a(p, k), aqk = a(q, k);
                    a(p, k) = c * apk - s * aqk;
                    a(q, k) = s * apk + c * aqk;
                }
                for (int k = 0; k < n; k++) {
                    const double vkp = V[(size_t)k * n + p], vkq = V[(size_t)k * n + q];
                    V[(size_t)k * n + p] = c * vkp - s * vkq;
                    V[(size_t)k * n + q] = s * vkp + c * vkq;
                }
            }
    }
    std::vector<int> idx(n);
    for (int i = 0; i < n; i++) idx[i] = i;
    std::sort(idx.begin(), idx.end(), [&](int x, int y) { return a(x, x) < a(y, y); });
    w.resize(n);
    std::vector<double> V2((size_t)n * n);
    for (int k = 0; k < n; k++) {
        w[k] = a(idx[k], idx[k]);
        for (int i = 0; i < n; i++) V2[(size_t)i * n + k] = V[(size_t)i * n + idx[k]];
    }
    V.swap(V2);
}

namespace {

double norm(const std::vector<double> &v)
{
    double s = 0;
    for (double x : v) s += x * x;
    return std::sqrt(s);
}

double dot(const std::vector<double> &a, const std::vector<double> &b)
{
    double s = 0;
    for (size_t i = 0; i < a.size(); i++) s += a[i] * b[i];
    return s;
}

} // namespace

NewtonDriver::NewtonDriver(const std::vector<int> &order, const std::vector<double> &scale, int nparams,
                           EvalFns fns)
    : order_(order), scale_(scale), P_(nparams), M_((int)order.size()), f_(std::move(fns))
{
}

std::vector<double> NewtonDriver::decompress(const std::vector<double> &step) const
{
    std::vector<double> out(M_);
    for (int k = 0; k < M_; k++) out[k] = scale_[k] * step[order_[k]];
    return out;
}

void NewtonDriver::record(double E, double gnorm)
{
    const double t = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
    trace.push_back({t, E, gnorm, hessianEvals, energyEvals});
}

// runNewtonMethod: avoidNegativeHessianValues = true (as optimise() calls it)
bool NewtonDriver::newton(std::vector<double> &angles)
{
    const double zeroThreshold = 1e-10;
    for (int count = 0; count < maxNewtonSteps; count++) {
        std::vector<double> H((size_t)P_ * P_), g(P_);
        double E = 0;
        if (!f_.hessian(angles, H, g, E)) return false;
        hessianEvals++;
        energyEvals++;
        record(E, norm(g));
        if (stopEnergy_ && E <= stopEnergyValue_) return true;
        std::vector<double> w, V;
        symmetric_eigen(P_, H, w, V);
        std::vector<double> step(P_, 0.0);
        for (int i = 0; i < P_; i++) {
            if (std::abs(w[i]) < zeroThreshold) continue;
            const double inv = std::abs(1.0 / w[i]);
            double proj = 0;
            for (int r = 0; r < P_; r++) proj += V[(size_t)r * P_ + i] * g[r];
            for (int r = 0; r < P_; r++) step[r] -= V[(size_t)r * P_ + i] * inv * proj;
        }
        const std::vector<double> upd = decompress(step);
        // cubic model along the step (src/tupsquantities.cpp:964-1122)
        std::vector<double> Hs(P_, 0.0);
        for (int i = 0; i < P_; i++)
            for (int j = 0; j < P_; j++) Hs[i] += H[(size_t)i * P_ + j] * step[j];
        long double tMax = 1.5, t = 1, newT = 1;
        const std::vector<double> anglesCopy = angles;
        auto place = [&](long double tt) {
            for (int k = 0; k < M_; k++) angles[k] = anglesCopy[k] + (double)(tt * upd[k]);
        };
        place(t);
        int searchCount = 0;
        const double gnorm = norm(g);
        while (true) {
            searchCount++;
            double Et = 0;
            std::vector<double> gt(P_);
            if (!f_.egrad(angles, Et, gt)) return false;
            energyEvals++;
            long double b = dot(step, Hs);
            const long double c = dot(step, g);
            const long double found = dot(gt, step);
            const long double a = (found - (b + c)) / (t * t);
            b /= t;
            const long double disc = b * b - 4 * a * c;
            if (a >= 0) {
                newT = (-b + std::sqrt(disc)) / (2 * a);
            } else if (disc >= 0) {
                const long double t1 = (-b + std::sqrt(disc)) / (2 * a);
                const long double E1 = a * t1 * t1 * t1 / 3 + b * t1 * t1 / 2 + c * t1 + E;
                const long double E2 = a * tMax * tMax * tMax / 3 + b * tMax * tMax / 2 + c * tMax + E;
                newT = (E2 < E1) ? tMax : t1;
            } else {
                newT = tMax;
            }
            place(newT);
            double stepNorm = 0;
            for (int k = 0; k < M_; k++) stepNorm += (double)(newT * upd[k]) * (double)(newT * upd[k]);
            stepNorm = std::sqrt(stepNorm);
            if (Et - E > 1e-10 && gnorm > 1e-8 && stepNorm > 1e-8) {
                t = t / 2;
                tMax = tMax / 2;
                place(t);
                if (searchCount > 20) break;
            } else {
                if (newT > tMax) newT = tMax;
                place(newT);
                break;
            }
        }
        if (gnorm < gradTolerance) break;
    }
    return true;
}

// doStepsUntilHessianIsPositiveDefinite with doDerivativeSteps = false
bool NewtonDriver::escape(std::vector<double> &angles, bool &foundAny)
{
    const double stepSize = 0.01;
    foundAny = false;
    for (int count = 0; count < 10; count++) {
        std::vector<double> H((size_t)P_ * P_), g(P_);
        double E;
        if (!f_.hessian(angles, H, g, E)) return false;
        hessianEvals++;
        std::vector<double> w, V;
        symmetric_eigen(P_, H, w, V);
        std::vector<double> step(P_, 0.0);
        bool found = false;
        for (int i = 0; i < P_; i++)
            if (w[i] < -1e-7) {
                found = true;
                foundAny = true;
                for (int r = 0; r < P_; r++) step[r] += V[(size_t)r * P_ + i] * stepSize;
            }
        const std::vector<double> upd = decompress(step);
        for (int k = 0; k < M_; k++) angles[k] += upd[k];
        if (!found) break;
    }
    return true;
}

bool NewtonDriver::optimise(std::vector<double> &angles, double &E)
{
    t0_ = std::chrono::steady_clock::now();
    bool amStuck = false;
    for (int outer = 0; outer < maxOuter; outer++) {
        if (!newton(angles)) return false;
        if (stopEnergy_ && !trace.empty() && trace.back().energy <= stopEnergyValue_) break;
        bool found = false;
        if (!escape(angles, found)) return false;
        std::vector<double> g(P_);
        if (!f_.egrad(angles, E, g)) return false;
        energyEvals++;
        const double gn = norm(g);
        record(E, gn);
        if (gn < gradTolerance) break;
        if (!found && amStuck) break;
        amStuck = !found;
    }
    std::vector<double> g(P_);
    return f_.egrad(angles, E, g);
}

} // namespace ae
