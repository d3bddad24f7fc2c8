// The C ABI (include/ansatzsynth.h).  Argument checks, their order and the
// return codes follow the reference's src/AnsatzSynthInterface.cpp; every
// call on a ctx takes that ctx's lock.  No exception crosses the boundary.
#include "../../include/ansatzsynth.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <limits>
#include <new>
#include <string>
#include <vector>

#include "manager.hpp"
#include "runfiles.hpp"

using ae::log_msg;
using ae::Manager;

namespace {
bool traceInterfaceCalls = false;

template <class T>
std::string join(const T *p, int n)
{
    std::string s = "[";
    for (int i = 0; i < n; i++) {
        if (i) s += ", ";
        s += std::to_string(p[i]);
    }
    return s + "]";
}
} // namespace

extern "C" {

void setTraceInterfaceCalls(int val) { traceInterfaceCalls = val != 0; }

void *init(void)
{
    Manager *ptr = new (std::nothrow) Manager;
    if (traceInterfaceCalls) log_msg("init: new context %p", (void *)ptr);
    return ptr;
}

int cleanup(void *ctx)
{
    if (traceInterfaceCalls) log_msg("cleanup(%p)", ctx);
    if (ctx == nullptr) {
        log_msg("cleanup: the handle pointer is null, so the context (if any) cannot be released");
        return 1;
    }
    void **myCtx = (void **)ctx;
    if (*myCtx == nullptr) return 0;
    delete static_cast<Manager *>(*myCtx);
    *myCtx = nullptr;
    return 0;
}

int setExcitation(int nparams, const int *operators, const int *orderfile, void *ctx)
{
    return setExcitationScale(nparams, operators, orderfile, nullptr, ctx);
}

int setExcitationScale(int nparams, const int *operators, const int *orderfile, const double *scale, void *ctx)
{
    if (ctx == nullptr) {
        log_msg("the context is null");
        return 1;
    }
    if (nparams < 1) {
        log_msg("setExcitation: the operator count is below 1");
        return 2;
    }
    if (traceInterfaceCalls) {
        log_msg("setExcitationScale(nparams=%d, operators=[%s], order=[%s])", nparams,
                join(operators, 4 * nparams).c_str(), join(orderfile, nparams).c_str());
        if (scale) log_msg("  scales [%s]", join(scale, nparams).c_str());
    }
    Manager *m = static_cast<Manager *>(ctx);
    std::lock_guard<std::mutex> lock(m->lock);
    std::vector<std::array<int8_t, 4>> excs(nparams);
    for (int i = 0; i < nparams; i++)
        for (int j = 0; j < 4; j++) {
            if (operators[nparams * j + i] - 1 > std::numeric_limits<int8_t>::max()) {
                log_msg("setExcitation: orbital index %d is above 128", operators[nparams * j + i]);
                return 3;
            }
            excs[i][j] = (int8_t)(operators[nparams * j + i] - 1);  // Fortran order
        }
    if (!m->storeOperators(excs)) return 4;
    std::vector<std::pair<int, double>> dep;
    for (int i = 0; i < nparams; i++) {
        if (orderfile[i] > nparams) {
            log_msg("setExcitation: parameter index %d is above the operator count", orderfile[i]);
            return 5;
        }
        dep.push_back({orderfile[i] - 1, scale ? scale[i] : 1.0});
    }
    if (!m->storeParameterDependencies(dep)) return 6;
    return 0;
}

int setHamiltonian(int n, const int *iindexes, const int *jindexes, const double *coeffs, void *ctx)
{
    if (ctx == nullptr) {
        log_msg("the context is null");
        return 1;
    }
    if (traceInterfaceCalls)
        log_msg("setHamiltonian(n=%d, i=[%s], j=[%s], ...)", n, join(iindexes, n).c_str(),
                join(jindexes, n).c_str());
    Manager *m = static_cast<Manager *>(ctx);
    std::lock_guard<std::mutex> lock(m->lock);
    if (n < 1) {
        log_msg("the entry count is below 1");
        return 2;
    }
    std::vector<int> iv(iindexes, iindexes + n), jv(jindexes, jindexes + n);
    for (int &x : iv) x--;
    for (int &x : jv) x--;
    std::vector<double> cv(coeffs, coeffs + n);
    if (!m->storeHamiltonian(std::move(iv), std::move(jv), std::move(cv))) return 3;
    return 0;
}

int setInitialState(int numqubits, int n, const int *iindexes, const double *coeffs, void *ctx)
{
    if (ctx == nullptr) {
        log_msg("the context is null");
        return 1;
    }
    if (traceInterfaceCalls)
        log_msg("setInitialState(qubits=%d, n=%d, indexes=[%s], coeffs=[%s])", numqubits, n,
                join(iindexes, n).c_str(), join(coeffs, n).c_str());
    Manager *m = static_cast<Manager *>(ctx);
    std::lock_guard<std::mutex> lock(m->lock);
    if (n < 1) {
        log_msg("the entry count is below 1");
        return 2;
    }
    if (numqubits < 1) {
        log_msg("the qubit count is below 1");
        return 2;
    }
    std::vector<int> iv(iindexes, iindexes + n);
    for (int &x : iv) x--;
    std::vector<double> cv(coeffs, coeffs + n);
    if (!m->storeInitial(numqubits, iv, cv)) return 3;
    return 0;
}

int setInitialStateComplex(int numqubits, int n, const int *iindexes, const ANSATZ_DOUBLE_COMPLEX *coeffs, void *ctx)
{
    if (ctx == nullptr) {
        log_msg("the context is null");
        return 1;
    }
    if (traceInterfaceCalls) log_msg("setInitialStateComplex(qubits=%d, n=%d)", numqubits, n);
    Manager *m = static_cast<Manager *>(ctx);
    std::lock_guard<std::mutex> lock(m->lock);
    if (n < 1) {
        log_msg("the entry count is below 1");
        return 2;
    }
    if (numqubits < 1) {
        log_msg("the qubit count is below 1");
        return 2;
    }
    if (iindexes == nullptr || coeffs == nullptr) {
        log_msg("the index or coefficient array is null");
        return 2;
    }
    std::vector<int> iv(iindexes, iindexes + n);
    for (int &x : iv) x--;
    std::vector<double> re(n), im(n);
    for (int k = 0; k < n; k++) {
        re[k] = coeffs[k].real();
        im[k] = coeffs[k].imag();
    }
    if (!m->storeInitialComplex(numqubits, iv, re, im)) return 3;
    return 0;
}

int getEnergy(int nangles, const double *angles, double *energy, void *ctx)
{
    if (ctx == nullptr) {
        log_msg("the context is null");
        return 1;
    }
    if (angles == nullptr) {
        log_msg("the angle array is null");
        return 2;
    }
    if (nangles < 1) {
        log_msg("the angle count is below 1");
        return 3;
    }
    if (traceInterfaceCalls) log_msg("getEnergy(nangles=%d)", nangles);
    Manager *m = static_cast<Manager *>(ctx);
    std::lock_guard<std::mutex> lock(m->lock);
    double E = 0;
    if (!m->getExpectationValue(std::vector<double>(angles, angles + nangles), E)) return 3;
    *energy = E;
    return 0;
}

int getFinalState(int nangles, const double *angles, int nbasisvectors, double *finalstate, void *ctx)
{
    if (ctx == nullptr) {
        log_msg("the context is null");
        return 1;
    }
    if (angles == nullptr) {
        log_msg("the angle array is null");
        return 2;
    }
    if (finalstate == nullptr) {
        log_msg("the final-state output array is null");
        return 3;
    }
    if (nbasisvectors < 1) {
        log_msg("getFinalState: the output length is below 1");
        return 4;
    }
    if (nangles < 1) {
        log_msg("the angle count is below 1");
        return 5;
    }
    if (traceInterfaceCalls) log_msg("getFinalState(nangles=%d, nbasis=%d)", nangles, nbasisvectors);
    Manager *m = static_cast<Manager *>(ctx);
    std::lock_guard<std::mutex> lock(m->lock);
    std::vector<double> state;
    if (!m->getFinalState(std::vector<double>(angles, angles + nangles), state)) return 6;
    if (state.size() != (size_t)nbasisvectors) {
        log_msg("getFinalState: the output holds %d amplitudes, the full space has %zu", nbasisvectors, state.size());
        return 7;
    }
    std::memcpy(finalstate, state.data(), state.size() * sizeof(double));
    return 0;
}

int getFinalStateComplex(int nangles, const double *angles, int nbasisvectors, ANSATZ_DOUBLE_COMPLEX *finalstate,
                         void *ctx)
{
    if (ctx == nullptr) {
        log_msg("the context is null");
        return 1;
    }
    if (angles == nullptr) {
        log_msg("the angle array is null");
        return 2;
    }
    if (finalstate == nullptr) {
        log_msg("the final-state output array is null");
        return 3;
    }
    if (nbasisvectors < 1) {
        log_msg("getFinalStateComplex: the output length is below 1");
        return 4;
    }
    if (nangles < 1) {
        log_msg("the angle count is below 1");
        return 5;
    }
    if (traceInterfaceCalls) log_msg("getFinalStateComplex(nangles=%d, nbasis=%d)", nangles, nbasisvectors);
    Manager *m = static_cast<Manager *>(ctx);
    std::lock_guard<std::mutex> lock(m->lock);
    std::vector<double> re, im;
    if (!m->getFinalStateComplex(std::vector<double>(angles, angles + nangles), re, im)) return 6;
    if (re.size() != (size_t)nbasisvectors) {
        log_msg("getFinalStateComplex: the output holds %d amplitudes, the full space has %zu", nbasisvectors,
                re.size());
        return 7;
    }
    for (size_t i = 0; i < re.size(); i++) finalstate[i] = ANSATZ_DOUBLE_COMPLEX(re[i], im[i]);
    return 0;
}

int getGradient_COMP(int nangles, const double *angles, double *gradient, void *ctx)
{
    if (ctx == nullptr) {
        log_msg("the context is null");
        return 1;
    }
    if (angles == nullptr) {
        log_msg("the angle array is null");
        return 2;
    }
    if (gradient == nullptr) {
        log_msg("the gradient output array is null");
        return 3;
    }
    if (nangles < 1) {
        log_msg("the angle count is below 1");
        return 4;
    }
    if (traceInterfaceCalls) log_msg("getGradient_COMP(nangles=%d)", nangles);
    Manager *m = static_cast<Manager *>(ctx);
    std::lock_guard<std::mutex> lock(m->lock);
    std::vector<double> g;
    if (!m->getGradientComp(std::vector<double>(angles, angles + nangles), g)) return 5;
    if ((size_t)nangles != g.size()) {
        log_msg("internal error: the gradient came back with the wrong length");
        return -6;
    }
    std::memcpy(gradient, g.data(), g.size() * sizeof(double));
    return 0;
}

int getGradientOperators(int nangles, const double *angles, double *gradient, void *ctx)
{
    if (ctx == nullptr) {
        log_msg("the context is null");
        return 1;
    }
    if (angles == nullptr) {
        log_msg("the angle array is null");
        return 2;
    }
    if (gradient == nullptr) {
        log_msg("the gradient output array is null");
        return 3;
    }
    if (nangles < 1) {
        log_msg("the angle count is below 1");
        return 4;
    }
    if (traceInterfaceCalls) log_msg("getGradientOperators(nangles=%d)", nangles);
    Manager *m = static_cast<Manager *>(ctx);
    std::lock_guard<std::mutex> lock(m->lock);
    std::vector<double> g;
    if (!m->getGradient(std::vector<double>(angles, angles + nangles), g)) return 5;
    std::memcpy(gradient, g.data(), g.size() * sizeof(double));
    return 0;
}

int getOverlapGradient(int nangles, const double *angles, int nbasisvectors, const double *projvector,
                       double *overlap, double *gradient, void *ctx)
{
    if (ctx == nullptr) {
        log_msg("the context is null");
        return 1;
    }
    if (angles == nullptr || projvector == nullptr) {
        log_msg("the angle or projection array is null");
        return 2;
    }
    if (gradient == nullptr) {
        log_msg("the gradient output array is null");
        return 3;
    }
    if (nangles < 1 || nbasisvectors < 1) {
        log_msg("the angle count or the basis size is below 1");
        return 4;
    }
    if (traceInterfaceCalls) log_msg("getOverlapGradient(nangles=%d)", nangles);
    Manager *m = static_cast<Manager *>(ctx);
    std::lock_guard<std::mutex> lock(m->lock);
    long long dim = 0;
    int nops = 0, nparams = 0;
    if (!m->getDimensions(dim, nops, nparams)) return 5;
    if (nangles != nparams) {
        log_msg("gradient length does not match the parameter count");
        return -6;
    }
    std::vector<double> g;
    double ov = 0;
    const std::vector<double> proj(projvector, projvector + nbasisvectors);
    if (!m->getOverlapGradient(std::vector<double>(angles, angles + nangles), proj, ov, g)) return 5;
    std::memcpy(gradient, g.data(), g.size() * sizeof(double));
    if (overlap) *overlap = ov;
    return 0;
}

int getHessianProj(int nangles, const double *angles, int nbasisvectors, const double *projvector,
                   double *hessian, double *gradient, double *overlap, void *ctx)
{
    if (ctx == nullptr) {
        log_msg("the context is null");
        return 1;
    }
    if (angles == nullptr || projvector == nullptr) {
        log_msg("the angle or projection array is null");
        return 2;
    }
    if (hessian == nullptr) {
        log_msg("the Hessian output array is null");
        return 3;
    }
    if (nangles < 1 || nbasisvectors < 1) {
        log_msg("the angle count or the basis size is below 1");
        return 4;
    }
    if (traceInterfaceCalls) log_msg("getHessianProj(nangles=%d)", nangles);
    Manager *m = static_cast<Manager *>(ctx);
    std::lock_guard<std::mutex> lock(m->lock);
    long long dim = 0;
    int nops = 0, nparams = 0;
    if (!m->getDimensions(dim, nops, nparams)) return 5;
    if (nangles != nparams) {
        log_msg("gradient length does not match the parameter count");
        return -6;
    }
    std::vector<double> H, g;
    double ov = 0;
    const std::vector<double> proj(projvector, projvector + nbasisvectors);
    if (!m->getHessianProj(std::vector<double>(angles, angles + nangles), proj, H, g, ov)) return 5;
    // row-major P x P from the engine; the interface is column-major (symmetric anyway)
    for (int i = 0; i < nparams; i++)
        for (int j = 0; j < nparams; j++) hessian[(size_t)nparams * j + i] = H[(size_t)i * nparams + j];
    if (gradient) std::memcpy(gradient, g.data(), g.size() * sizeof(double));
    if (overlap) *overlap = ov;
    return 0;
}

int getHessian_COMP(int nangles, const double *angles, double *hessian, void *ctx)
{
    if (ctx == nullptr) {
        log_msg("the context is null");
        return 1;
    }
    if (angles == nullptr) {
        log_msg("the angle array is null");
        return 2;
    }
    if (hessian == nullptr) {
        log_msg("the Hessian output array is null");
        return 3;
    }
    if (nangles < 1) {
        log_msg("the angle count is below 1");
        return 4;
    }
    if (traceInterfaceCalls) log_msg("getHessian_COMP(nangles=%d)", nangles);
    Manager *m = static_cast<Manager *>(ctx);
    std::lock_guard<std::mutex> lock(m->lock);
    std::vector<double> h;
    if (!m->getHessianComp(std::vector<double>(angles, angles + nangles), h)) return 5;
    const int P = m->numberOfParameters();
    if (nangles != P) {
        log_msg("internal error: the Hessian came back with the wrong size");
        return -6;
    }
    for (int i = 0; i < P; i++)
        for (int j = 0; j < P; j++) hessian[(size_t)P * j + i] = h[(size_t)i * P + j];
    return 0;
}

// ------------------------------------------------------------ extensions
int setHamiltonianIntegrals(int norb, const double *h, const double *g, void *ctx)
{
    if (ctx == nullptr) {
        log_msg("the context is null");
        return 1;
    }
    if (norb < 1 || h == nullptr || g == nullptr) {
        log_msg("norb < 1 or integrals missing");
        return 2;
    }
    if (traceInterfaceCalls) log_msg("setHamiltonianIntegrals(norb=%d)", norb);
    Manager *m = static_cast<Manager *>(ctx);
    std::lock_guard<std::mutex> lock(m->lock);
    if (!m->storeIntegrals(norb, h, g)) return 3;
    return 0;
}

int setNuclearEnergy(double energy, void *ctx)
{
    if (ctx == nullptr) {
        log_msg("the context is null");
        return 1;
    }
    Manager *m = static_cast<Manager *>(ctx);
    std::lock_guard<std::mutex> lock(m->lock);
    if (!m->storeNuclearEnergy(energy)) return 3;
    return 0;
}

int setRunPath(const char *runpath, void *ctx)
{
    if (ctx == nullptr) {
        log_msg("the context is null");
        return 1;
    }
    if (runpath == nullptr || runpath[0] == 0) {
        log_msg("setRunPath: empty path");
        return 2;
    }
    if (traceInterfaceCalls) log_msg("setRunPath(%s)", runpath);
    Manager *m = static_cast<Manager *>(ctx);
    std::lock_guard<std::mutex> lock(m->lock);
    if (!m->storeRunPath(runpath)) return 3;
    return 0;
}

int loadRunFiles(const char *prefix, int withnuclear, void *ctx)
{
    if (ctx == nullptr) {
        log_msg("the context is null");
        return 1;
    }
    if (prefix == nullptr || prefix[0] == 0) {
        log_msg("setRunPath: empty path");
        return 2;
    }
    if (traceInterfaceCalls) log_msg("loadRunFiles called with %s, withnuclear %d", prefix, withnuclear);
    const std::string pre(prefix);
    int nq = 0;
    std::vector<int> idx;
    std::vector<double> coeff;
    std::vector<std::array<int8_t, 4>> ops;
    std::vector<std::pair<int, double>> dep;
    double nuc = 0;
    bool haveNuc = false;
    std::string err;
    if (!ae::files::read_initial(pre + "_Initial.dat", nq, idx, coeff, err) ||
        !ae::files::read_operators(pre + "_Operators.dat", ops, err) ||
        !ae::files::read_order(pre + "_Order.dat", dep, err) ||
        (withnuclear && !ae::files::read_nuclear_energy(pre + "_Nuclear_Energy.dat", nuc, haveNuc, err))) {
        log_msg("loadRunFiles: %s", err.c_str());
        return 3;
    }
    if (dep.size() != ops.size()) {
        log_msg("loadRunFiles: %zu operators but %zu order entries", ops.size(), dep.size());
        return 3;
    }
    Manager *m = static_cast<Manager *>(ctx);
    std::lock_guard<std::mutex> lock(m->lock);
    if (!m->storeInitial(nq, idx, coeff) || !m->storeOperators(ops) || !m->storeParameterDependencies(dep) ||
        (withnuclear && !m->storeNuclearEnergy(nuc)) || !m->storeRunPath(pre))
        return 4;
    return 0;
}

int readParameterSets(const char *path, int maxsets, int maxparams, int *nsets, int *nparams, double *energies,
                      double *params)
{
    if (path == nullptr || nsets == nullptr || nparams == nullptr) {
        log_msg("readParameterSets: NULL argument");
        return 2;
    }
    std::vector<double> E;
    std::vector<std::vector<double>> sets;
    std::string err;
    if (!ae::files::read_parameter_sets(path, E, sets, err)) {
        log_msg("readParameterSets: %s", err.c_str());
        return 3;
    }
    int np = 0;
    for (const auto &v : sets) np = std::max(np, (int)v.size());
    *nsets = (int)sets.size();
    *nparams = np;
    if (maxsets <= 0) return 0;  // counts only
    if ((int)sets.size() > maxsets || np > maxparams || energies == nullptr || params == nullptr) return 6;
    for (size_t s = 0; s < sets.size(); s++) {
        energies[s] = E[s];
        for (int i = 0; i < maxparams; i++) params[s * maxparams + i] = i < (int)sets[s].size() ? sets[s][i] : 0.0;
    }
    return 0;
}

int setHamiltonianStorage(int mode, void *ctx)
{
    if (ctx == nullptr) {
        log_msg("the context is null");
        return 1;
    }
    if (mode < 0 || mode > 2) {
        log_msg("storage mode must be 0, 1 or 2");
        return 2;
    }
    Manager *m = static_cast<Manager *>(ctx);
    std::lock_guard<std::mutex> lock(m->lock);
    if (!m->storeHamiltonianStorage(mode)) return 3;
    return 0;
}

int getEnergyGradient(int nangles, const double *angles, double *energy, double *gradient, void *ctx)
{
    if (ctx == nullptr) {
        log_msg("the context is null");
        return 1;
    }
    if (angles == nullptr) {
        log_msg("the angle array is null");
        return 2;
    }
    if (gradient == nullptr) {
        log_msg("the gradient output array is null");
        return 3;
    }
    if (nangles < 1) {
        log_msg("the angle count is below 1");
        return 4;
    }
    Manager *m = static_cast<Manager *>(ctx);
    std::lock_guard<std::mutex> lock(m->lock);
    std::vector<double> g;
    double E = 0;
    if (!m->getEnergyGradient(std::vector<double>(angles, angles + nangles), E, g)) return 5;
    if ((size_t)nangles != g.size()) {
        log_msg("Gradient is the wrong size: nangles must equal the number of parameters");
        return -6;
    }
    std::memcpy(gradient, g.data(), g.size() * sizeof(double));
    if (energy) *energy = E;
    return 0;
}

int getEnergies(int nvec, int nangles, const double *angles, double *energies, void *ctx)
{
    if (ctx == nullptr) {
        log_msg("the context is null");
        return 1;
    }
    if (angles == nullptr || energies == nullptr) {
        log_msg("nullptr angles or energies passed");
        return 2;
    }
    if (nvec < 1 || nangles < 1) {
        log_msg("nvec < 1 or NAngles < 1");
        return 3;
    }
    Manager *m = static_cast<Manager *>(ctx);
    std::lock_guard<std::mutex> lock(m->lock);
    std::vector<double> E;
    if (!m->getExpectationValues(nvec, nangles, angles, E)) return 5;
    std::memcpy(energies, E.data(), sizeof(double) * nvec);
    return 0;
}

int getEnergyGradients(int nvec, int nangles, const double *angles, double *energies, double *gradients, void *ctx)
{
    if (ctx == nullptr) {
        log_msg("the context is null");
        return 1;
    }
    if (angles == nullptr || energies == nullptr || gradients == nullptr) {
        log_msg("nullptr angles, energies or gradients passed");
        return 2;
    }
    if (nvec < 1 || nangles < 1) {
        log_msg("nvec < 1 or NAngles < 1");
        return 3;
    }
    Manager *m = static_cast<Manager *>(ctx);
    std::lock_guard<std::mutex> lock(m->lock);
    std::vector<double> E, g;
    if (!m->getEnergyGradients(nvec, nangles, angles, E, g)) return 5;
    if (g.size() != (size_t)nvec * nangles) {
        log_msg("Gradient is the wrong size: nangles must equal the number of parameters");
        return -6;
    }
    std::memcpy(energies, E.data(), sizeof(double) * nvec);
    std::memcpy(gradients, g.data(), sizeof(double) * g.size());
    return 0;
}

int getDimensions(long long *sectordim, int *noperators, int *nparams, void *ctx)
{
    if (ctx == nullptr) {
        log_msg("the context is null");
        return 1;
    }
    Manager *m = static_cast<Manager *>(ctx);
    std::lock_guard<std::mutex> lock(m->lock);
    long long d;
    int no, np;
    if (!m->getDimensions(d, no, np)) return 5;
    if (sectordim) *sectordim = d;
    if (noperators) *noperators = no;
    if (nparams) *nparams = np;
    return 0;
}

int setEngineOption(int key, int value, void *ctx)
{
    if (ctx == nullptr) {
        log_msg("the context is null");
        return 1;
    }
    Manager *m = static_cast<Manager *>(ctx);
    std::lock_guard<std::mutex> lock(m->lock);
    return m->setEngineOption(key, value) ? 0 : 5;
}

int getEngineOption(int key, int *value, void *ctx)
{
    if (ctx == nullptr) {
        log_msg("the context is null");
        return 1;
    }
    if (value == nullptr) {
        log_msg("nullptr value passed");
        return 2;
    }
    Manager *m = static_cast<Manager *>(ctx);
    std::lock_guard<std::mutex> lock(m->lock);
    int v = 0;
    if (!m->getEngineOption(key, v)) return 5;
    *value = v;
    return 0;
}

int autotune(int nvec, int nangles, const double *angles, const char *cachefile, int reps, char *report,
             int reportlen, void *ctx)
{
    if (ctx == nullptr) {
        log_msg("the context is null");
        return 1;
    }
    if (angles == nullptr) {
        log_msg("the angle array is null");
        return 2;
    }
    if (nvec < 1 || nangles < 1) {
        log_msg("nvec or nangles < 1");
        return 3;
    }
    Manager *m = static_cast<Manager *>(ctx);
    std::lock_guard<std::mutex> lock(m->lock);
    std::string rep;
    if (!m->autotune(nvec, nangles, angles, cachefile, reps, rep)) return 5;
    if (report && reportlen > 0) {
        const size_t n = std::min<size_t>(rep.size(), (size_t)reportlen - 1);
        std::memcpy(report, rep.data(), n);
        report[n] = '\0';
    }
    return 0;
}

} // extern "C"

extern "C" {

int uploadAngleBatch(int nvec, int nangles, const double *angles, void *ctx)
{
    if (ctx == nullptr) {
        log_msg("the context is null");
        return 1;
    }
    if (angles == nullptr) {
        log_msg("the angle array is null");
        return 2;
    }
    if (nvec < 1 || nangles < 1) {
        log_msg("nvec < 1 or NAngles < 1");
        return 3;
    }
    Manager *m = static_cast<Manager *>(ctx);
    std::lock_guard<std::mutex> lock(m->lock);
    return m->uploadAngleBatch(nvec, nangles, angles) ? 0 : 5;
}

int runAngleBatch(int withgradient, void *ctx)
{
    if (ctx == nullptr) {
        log_msg("the context is null");
        return 1;
    }
    Manager *m = static_cast<Manager *>(ctx);
    std::lock_guard<std::mutex> lock(m->lock);
    return m->runAngleBatch(withgradient != 0) ? 0 : 5;
}

int downloadAngleBatch(double *energies, double *gradients, void *ctx)
{
    if (ctx == nullptr) {
        log_msg("the context is null");
        return 1;
    }
    if (energies == nullptr) {
        log_msg("nullptr energies passed");
        return 2;
    }
    Manager *m = static_cast<Manager *>(ctx);
    std::lock_guard<std::mutex> lock(m->lock);
    if (gradients && m->batchAngleCount() != m->numberOfParameters()) {
        log_msg("Gradients need per-parameter angles in the uploaded batch");
        return -6;
    }
    return m->downloadAngleBatch(energies, gradients) ? 0 : 5;
}

int setKernelTiming(int on, void *ctx)
{
    if (ctx == nullptr) {
        log_msg("the context is null");
        return 1;
    }
    Manager *m = static_cast<Manager *>(ctx);
    std::lock_guard<std::mutex> lock(m->lock);
    return m->setKernelTiming(on != 0) ? 0 : 5;
}

int getKernelTiming(int kernelclass, double *milliseconds, long long *launches, void *ctx)
{
    if (ctx == nullptr) {
        log_msg("the context is null");
        return 1;
    }
    if (kernelclass < 0 || kernelclass > 3 || milliseconds == nullptr || launches == nullptr) {
        log_msg("bad kernel class or NULL output");
        return 2;
    }
    Manager *m = static_cast<Manager *>(ctx);
    std::lock_guard<std::mutex> lock(m->lock);
    double ms = 0;
    long long n = 0;
    if (!m->getKernelTiming(kernelclass, ms, n)) return 5;
    *milliseconds = ms;
    *launches = n;
    return 0;
}

} // extern "C"

extern "C" int optimise(int nangles, double *angles, double *energy, double stopenergy, int maxtrace, double *trace,
                        int *ntrace, void *ctx)
{
    if (ctx == nullptr) {
        log_msg("the context is null");
        return 1;
    }
    if (angles == nullptr || energy == nullptr) {
        log_msg("nullptr angles or energy passed");
        return 2;
    }
    if (nangles < 1) {
        log_msg("the angle count is below 1");
        return 3;
    }
    Manager *m = static_cast<Manager *>(ctx);
    std::lock_guard<std::mutex> lock(m->lock);
    std::vector<double> a(angles, angles + nangles), tr;
    double E = 0;
    const bool useStop = !std::isnan(stopenergy);
    if (!m->optimise(a, E, tr, stopenergy, useStop)) return 5;
    std::memcpy(angles, a.data(), sizeof(double) * nangles);
    *energy = E;
    if (trace && ntrace) {
        const int rows = (int)(tr.size() / 5);
        const int w = rows < maxtrace ? rows : maxtrace;
        std::memcpy(trace, tr.data(), sizeof(double) * 5 * (size_t)(w > 0 ? w : 0));
        *ntrace = w;
    }
    return 0;
}

extern "C" int iterativeOptimise(int nangles, double *angles, double *energy, int ediis, int maxtrace, double *trace,
                                 int *ntrace, void *ctx)
{
    if (ctx == nullptr) {
        log_msg("the context is null");
        return 1;
    }
    if (angles == nullptr || energy == nullptr) {
        log_msg("the angle or energy array is null");
        return 2;
    }
    if (nangles < 1) {
        log_msg("the angle count is below 1");
        return 3;
    }
    if (traceInterfaceCalls) log_msg("iterativeOptimise(nangles=%d, ediis=%d)", nangles, ediis);
    Manager *m = static_cast<Manager *>(ctx);
    std::lock_guard<std::mutex> lock(m->lock);
    std::vector<double> a(angles, angles + nangles), tr;
    double E = 0;
    if (!m->iterativeOptimise(a, E, tr, ediis != 0)) return 5;
    std::memcpy(angles, a.data(), sizeof(double) * nangles);
    *energy = E;
    if (trace && ntrace) {
        const int rows = (int)(tr.size() / 5);
        const int w = rows < maxtrace ? rows : maxtrace;
        std::memcpy(trace, tr.data(), sizeof(double) * 5 * (size_t)(w > 0 ? w : 0));
        *ntrace = w;
    }
    return 0;
}

extern "C" int subspaceDiagonalise(int nsets, int nangles, const double *angles, double *eigenvalues, double *hmat,
                                   double *smat, int *nullrank, void *ctx)
{
    if (ctx == nullptr) {
        log_msg("the context is null");
        return 1;
    }
    if (angles == nullptr || eigenvalues == nullptr) {
        log_msg("the angle or eigenvalue array is null");
        return 2;
    }
    if (nsets < 1 || nangles < 1) {
        log_msg("the set count or the angle count is below 1");
        return 3;
    }
    if (traceInterfaceCalls) log_msg("subspaceDiagonalise(nsets=%d, nangles=%d)", nsets, nangles);
    Manager *m = static_cast<Manager *>(ctx);
    std::lock_guard<std::mutex> lock(m->lock);
    std::vector<double> ev, H, S;
    int nr = 0;
    if (!m->subspaceDiagonalise(nsets, nangles, angles, ev, H, S, nr)) return 5;
    std::memcpy(eigenvalues, ev.data(), sizeof(double) * nsets);
    // symmetric: row-major = column-major
    if (hmat) std::memcpy(hmat, H.data(), sizeof(double) * H.size());
    if (smat) std::memcpy(smat, S.data(), sizeof(double) * S.size());
    if (nullrank) *nullrank = nr;
    return 0;
}

extern "C" int getProperties(int nsets, int nangles, const double *angles, int noverlaps, int nohess, double *props,
                             double *hessians, double *metrics, double *eigenvalues, void *ctx)
{
    if (ctx == nullptr) {
        log_msg("the context is null");
        return 1;
    }
    if (angles == nullptr || props == nullptr || (noverlaps > 0 && eigenvalues == nullptr)) {
        log_msg("the angle, property or eigenvalue array is null");
        return 2;
    }
    if (nsets < 1 || nangles < 1 || noverlaps < 0 || noverlaps > 16) {
        log_msg("the set count or angle count is below 1, or the overlap count outside 0 .. 16");
        return 3;
    }
    if (traceInterfaceCalls) log_msg("getProperties(nsets=%d, nangles=%d, noverlaps=%d)", nsets, nangles, noverlaps);
    Manager *m = static_cast<Manager *>(ctx);
    std::lock_guard<std::mutex> lock(m->lock);
    std::vector<double> rows, H, G, low;
    if (!m->properties(nsets, nangles, angles, noverlaps, nohess != 0, rows, H, G, low)) return 5;
    std::memcpy(props, rows.data(), sizeof(double) * rows.size());
    const int P = m->numberOfParameters();
    // per set P x P, symmetric: row-major = column-major
    if (!nohess && hessians) std::memcpy(hessians, H.data(), sizeof(double) * (size_t)nsets * P * P);
    if (!nohess && metrics) std::memcpy(metrics, G.data(), sizeof(double) * (size_t)nsets * P * P);
    if (noverlaps > 0) std::memcpy(eigenvalues, low.data(), sizeof(double) * noverlaps);
    return 0;
}

// One process-wide lock: every no-allocation call on a GPU shares that GPU's
// scratch array, so calls from this process run one at a time.
static std::mutex noAllocLock;

extern "C" int getEnergyGradientNoAlloc(int nangles, const double *angles, double *energy, double *gradient,
                                        void *ctx)
{
    if (ctx == nullptr) {
        log_msg("the context is null");
        return 1;
    }
    if (angles == nullptr) {
        log_msg("the angle array is null");
        return 2;
    }
    if (gradient == nullptr) {
        log_msg("the gradient output array is null");
        return 3;
    }
    if (nangles < 1) {
        log_msg("the angle count is below 1");
        return 4;
    }
    Manager *m = static_cast<Manager *>(ctx);
    std::lock_guard<std::mutex> lock(m->lock);
    if (!m->built()) {
        log_msg("getEnergyGradientNoAlloc: the context is not built yet (an ordinary evaluation builds it)");
        return 6;
    }
    if (nangles != m->numberOfParameters()) {
        log_msg("gradient length does not match the parameter count");
        return -6;
    }
    std::lock_guard<std::mutex> glock(noAllocLock);
    double E = 0;
    const int rc = m->getEnergyGradientNoAlloc(angles, nangles, E, gradient);
    if (rc == 1) return 6;
    if (rc == 2) return 7;
    if (rc != 0) return 5;
    if (energy) *energy = E;
    return 0;
}
