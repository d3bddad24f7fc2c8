// The reference's iterative optimiser: an outer fixed-point loop that
// freezes the projection vector p = H psi(previous angles) and minimises the
// projected energy f(theta) = <p|psi(theta)> by Newton steps, restating
// TUPSQuantities::iterativeTups(FusedEvolve&) (src/tupsquantities.cpp:
// 2044-2145) and runNewtonMethodProjected (:1238-1465).  The fixed point is
// a stationary point of E: d/dtheta <H psi|psi(theta)> = (1/2) dE/dtheta.
//
// Optional energy-DIIS extrapolation of the outer iterates (src/diis.h;
// the reference compiles it but leaves it switched off, doDIIS = false at
// src/tupsquantities.cpp:2050): coefficients c_i >= 0, sum c_i = 1,
// minimising c^T B c with B_ij = <e_i|e_j>, e_i = psi_i - psi_(i-1); the
// constraint is built in by c_i = t_i^2 and a Lagrange multiplier, solved by
// damped Newton-Raphson on (t, lambda).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <deque>

#include "newton.hpp"

namespace ae {

namespace {

double vdot(const std::vector<double> &a, const std::vector<double> &b)
{
    double s = 0;
    for (size_t i = 0; i < a.size(); i++) s += a[i] * b[i];
    return s;
}

double vnorm(const std::vector<double> &a) { return std::sqrt(vdot(a, a)); }

// -V |lambda|^-1 V^T g with eigenvalues below `zero` in magnitude dropped
std::vector<double> spectral_step(int P, const std::vector<double> &H, const std::vector<double> &g, double zero)
{
    std::vector<double> w, V;
    symmetric_eigen(P, H, w, V);
    std::vector<double> y(P, 0.0), step(P, 0.0);
    for (int k = 0; k < P; k++) {
        if (std::fabs(w[k]) < zero) continue;
        double c = 0;
        for (int i = 0; i < P; i++) c += V[(size_t)i * P + k] * g[i];
        y[k] = c / std::fabs(w[k]);
    }
    for (int i = 0; i < P; i++) {
        double s = 0;
        for (int k = 0; k < P; k++) s += V[(size_t)i * P + k] * y[k];
        step[i] = -s;
    }
    return step;
}

} // namespace

ProjectedDriver::ProjectedDriver(const std::vector<int> &order, const std::vector<double> &scale, int nparams,
                                 ProjFns fns)
    : order_(order), scale_(scale), P_(nparams), M_((int)order.size()), f_(std::move(fns))
{
}

std::vector<double> ProjectedDriver::to_ops(const std::vector<double> &theta) const
{
    std::vector<double> x(M_);
    for (int k = 0; k < M_; k++) x[k] = scale_[k] * theta[order_[k]];
    return x;
}

// First operator of each parameter, divided by its scale (the reference's
// normCompress, src/tupsquantities.cpp:655-666).
std::vector<double> ProjectedDriver::to_params(const std::vector<double> &x) const
{
    std::vector<double> theta(P_, 0.0);
    std::vector<bool> seen(P_, false);
    for (int k = 0; k < M_; k++)
        if (!seen[order_[k]]) {
            theta[order_[k]] = x[k] / scale_[k];
            seen[order_[k]] = true;
        }
    return theta;
}

// Newton on f(theta) = <p|psi(theta)> at a fixed p, at most maxNewtonSteps
// steps, until |g| < 1e-10.  Step -V|lambda|^-1 V^T g (|lambda| < 1e-10
// dropped), then a search along it on the cubic model
// f(t) = a t^3 / 3 + b t^2 / 2 + c t + f0 fitted to f0, the gradient at 0,
// the Hessian's curvature and the directional derivative at the trial
// point; trial steps double while they keep lowering f (a forward run) and
// halve while they raise it, at most 5 model fits
// (src/tupsquantities.cpp:1317-1441).
bool ProjectedDriver::newton(std::vector<double> &x)
{
    std::vector<double> H, g, gt;
    for (int it = 0; it < maxNewtonSteps; it++) {
        double f0 = 0;
        if (!f_.hessianProj(x, H, g, f0)) return false;
        hessianEvals++;
        energyEvals++;
        const double gn = vnorm(g);
        const std::vector<double> u = spectral_step(P_, H, g, 1e-10);  // per parameter
        const std::vector<double> du = to_ops(u);
        double curv = 0;  // u^T H u
        for (int i = 0; i < P_; i++) {
            double r = 0;
            for (int j = 0; j < P_; j++) r += H[(size_t)i * P_ + j] * u[j];
            curv += u[i] * r;
        }
        const double slope = vdot(u, g);
        const std::vector<double> x0 = x;
        auto at = [&](long double t) {
            for (int k = 0; k < M_; k++) x[k] = x0[k] + (double)(t * du[k]);
        };
        long double t = 1, tMax = 1.5, tNext = 1, lastForward = 1;
        bool forward = false;
        at(t);
        for (int fits = 1;; fits++) {
            double ft = 0;
            if (!f_.gradProj(x, ft, gt)) return false;
            energyEvals++;
            const long double dT = vdot(gt, u);
            long double a = (dT - ((long double)curv + slope)) / (t * t);
            const long double b = (long double)curv / t, c = slope;
            if (ft < f0 && gn > 1e-5) forward = true;
            const long double disc = b * b - 4 * a * c;
            auto model = [&](long double s) { return a * s * s * s / 3 + b * s * s / 2 + c * s + f0; };
            if (a >= 0) {
                tNext = (-b + std::sqrt(disc)) / (2 * a);
            } else if (disc >= 0) {
                const long double t1 = (-b + std::sqrt(disc)) / (2 * a);
                tNext = model(tMax) < model(t1) ? tMax : t1;
            } else {
                tNext = tMax;  // the model falls all the way to the bound
            }
            tNext = std::min(tNext, tMax);
            at(tNext);
            if (fits > 5) break;
            if (ft > f0 && gn > 1e-5) {  // the trial point went up
                if (forward) {
                    lastForward = std::min(lastForward, tMax / 2);
                    tMax /= 2;
                    at(lastForward);
                    break;
                }
                t /= 2;
                tMax /= 2;
                at(t);
            } else if (forward) {  // still going down: try twice as far
                t *= 2;
                tMax *= 2;
                at(t);
                lastForward = tNext;
            } else {
                at(std::min(tNext, tMax));
                break;
            }
        }
        if (gn < 1e-10) break;
    }
    return true;
}

bool ProjectedDriver::optimise(std::vector<double> &x, double &E)
{
    t0_ = std::chrono::steady_clock::now();
    x = to_ops(to_params(x));  // onto the parameter space, as the reference does on entry
    std::vector<double> prev, dest, g;
    if (!f_.setProjection(x, prev)) return false;  // p = H psi(x); prev = psi(x)
    std::deque<std::vector<double>> hisTheta, hisErr;
    std::vector<double> xNewton;  // the last Newton result, whose energy E is
    for (int outer = 0; outer < maxOuter; outer++) {
        if (!newton(x)) return false;
        if (!f_.setProjection(x, dest)) return false;
        double f = 0;
        if (!f_.gradProj(x, f, g)) return false;
        energyEvals++;
        E = f;  // <H psi|psi> at p = H psi(x)
        xNewton = x;
        const double gn = vnorm(g);
        trace.push_back({std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count(), E, gn,
                         hessianEvals, energyEvals});
        if (gn < gradTolerance) break;
        if (useEdiis) {
            std::vector<double> err(dest.size());
            for (size_t i = 0; i < dest.size(); i++) err[i] = dest[i] - prev[i];
            hisTheta.push_back(to_params(x));
            hisErr.push_back(std::move(err));
            if ((int)hisTheta.size() > ediisSize) {
                hisTheta.pop_front();
                hisErr.pop_front();
            }
            const int n = (int)hisTheta.size();
            std::vector<double> B((size_t)n * n);
            for (int i = 0; i < n; i++)
                for (int j = 0; j < n; j++) B[(size_t)i * n + j] = vdot(hisErr[i], hisErr[j]);
            const std::vector<double> c = ediis_coefficients(n, B);
            std::vector<double> theta(P_, 0.0);
            for (int i = 0; i < n; i++)
                for (int p = 0; p < P_; p++) theta[p] += c[i] * hisTheta[i][p];
            // the projection stays H psi of the last Newton result, as in the reference
            x = to_ops(theta);
        }
        prev.swap(dest);
    }
    // a run that ends on the iteration limit after an extrapolation returns
    // the last Newton result, the point its energy belongs to (the
    // reference's loop, whose DIIS branch is switched off, would return the
    // extrapolated angles with the earlier energy)
    if (!xNewton.empty()) x = xNewton;
    return true;
}

// Coefficients c_i = t_i^2 minimising c^T B c with sum c_i = 1: Newton-Raphson
// on the stationarity conditions of L(t, lambda) = (t^2)^T B t^2 -
// lambda (sum t_i^2 - 1), from t_i = sqrt(n), lambda = 1, at most 200 steps;
// eigenvalues of the Jacobian below 1e-6 in magnitude are dropped, and a step
// that increases the residual is halved until it does not (src/diis.h,
// src/diis.cpp).
std::vector<double> ediis_coefficients(int n, const std::vector<double> &B)
{
    const int m = n + 1;
    std::vector<double> z(m, std::sqrt((double)n));
    z[n] = 1.0;
    auto residual = [&](const std::vector<double> &p) {
        std::vector<double> r(m, 0.0);
        double sq = 0;
        for (int i = 0; i < n; i++) {
            double bt = 0;
            for (int j = 0; j < n; j++) bt += B[(size_t)i * n + j] * p[j] * p[j];
            r[i] = 4 * bt * p[i] - 2 * p[n] * p[i];
            sq += p[i] * p[i];
        }
        r[n] = 1 - sq;
        return r;
    };
    for (int it = 0; it < 200; it++) {
        const std::vector<double> r = residual(z);
        const double rn = vnorm(r);
        if (rn < 1e-12) break;
        std::vector<double> J((size_t)m * m, 0.0);
        for (int i = 0; i < n; i++) {
            double bt = 0;
            for (int k = 0; k < n; k++) bt += B[(size_t)i * n + k] * z[k] * z[k];
            for (int j = 0; j < n; j++)
                J[(size_t)i * m + j] = 8 * z[i] * B[(size_t)i * n + j] * z[j] + (i == j ? 4 * bt - 2 * z[n] : 0.0);
            J[(size_t)i * m + n] = J[(size_t)n * m + i] = -2 * z[i];
        }
        std::vector<double> w, V;
        symmetric_eigen(m, J, w, V);
        std::vector<double> step(m, 0.0);
        for (int k = 0; k < m; k++) {
            if (std::fabs(w[k]) < 1e-6) continue;
            double c = 0;
            for (int i = 0; i < m; i++) c += V[(size_t)i * m + k] * r[i];
            for (int i = 0; i < m; i++) step[i] -= V[(size_t)i * m + k] * c / w[k];
        }
        for (int i = 0; i < m; i++) z[i] += step[i];
        double smax = 0;
        for (double s : step) smax = std::max(smax, std::fabs(s));
        while (vnorm(residual(z)) > rn && smax > 1e-15) {
            for (int i = 0; i < m; i++) {
                step[i] /= 2;
                z[i] -= step[i];
            }
            smax /= 2;
        }
        if (!(smax > 1e-15)) break;
    }
    std::vector<double> c(n);
    for (int i = 0; i < n; i++) c[i] = z[i] * z[i];
    return c;
}

} // namespace ae
