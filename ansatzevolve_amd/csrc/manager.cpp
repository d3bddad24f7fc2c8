// Host-side counterpart of stateAnsatzManager (src/AnsatzManager.cpp): the
// stored set-up, the sector choice, the lazily built device engine and the
// angle conversions.
#include "manager.hpp"
#include "runfiles.hpp"
#include "spectral.hpp"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <map>

namespace ae {

void log_msg(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    std::fputs("[ansatzsynth] ", stderr);
    std::vfprintf(stderr, fmt, ap);
    std::fputc('\n', stderr);
    va_end(ap);
}

bool Manager::frozen(const char *what) const
{
    if (!eng_) return false;
    log_msg("%s rejected: the set-up is fixed once the first evaluation has run", what);
    return true;
}

// ---------------------------------------------------------------- set-up

bool Manager::storeOperators(const std::vector<std::array<int8_t, 4>> &excs)
{
    const bool locked = frozen("Operators");
    if (excs.empty()) log_msg("Operators rejected: the list is empty");
    if (locked || excs.empty()) return false;
    ops_ = excs;
    return true;
}

// The start state, and which occupations all of its nonzero components
// share (they decide the sector: src/AnsatzManager.cpp:271-374).  Bits
// 0 .. q/2-1 are spin down, q/2 .. q-1 spin up.
bool Manager::storeInitial(int numberOfQubits, const std::vector<int> &indexes, const std::vector<double> &coeffs)
{
    return storeStart(numberOfQubits, indexes, coeffs, nullptr);
}

// setInitialStateComplex (src/AnsatzSynthInterface.cpp:251-300): the same
// checks, amplitudes re + i im; the context then evaluates complex
// amplitudes.
bool Manager::storeInitialComplex(int numberOfQubits, const std::vector<int> &indexes, const std::vector<double> &re,
                                  const std::vector<double> &im)
{
    return storeStart(numberOfQubits, indexes, re, &im);
}

bool Manager::storeStart(int numberOfQubits, const std::vector<int> &indexes, const std::vector<double> &coeffs,
                         const std::vector<double> *imag)
{
    std::vector<std::string> why;
    if (frozen("Initial state")) why.push_back("set-up fixed");
    if (numberOfQubits < 1 || numberOfQubits > 30) {
        log_msg("Initial state rejected: %d qubits (1 to 30 are supported)", numberOfQubits);
        why.push_back("qubits");
    }
    if (indexes.empty() || coeffs.empty()) {
        log_msg("Initial state rejected: no components");
        why.push_back("empty");
    } else if (indexes.size() != coeffs.size()) {
        log_msg("Initial state rejected: %zu indexes against %zu coefficients", indexes.size(), coeffs.size());
        why.push_back("sizes");
    }
    if (why.empty())
        for (int b : indexes)
            if (b < 0 || (long long)b >= (1ll << numberOfQubits)) {
                log_msg("Initial state rejected: basis index %d is outside 1 .. 2^%d", b + 1, numberOfQubits);
                why.push_back("range");
                break;
            }
    if (!why.empty()) return false;

    auto im = [&](size_t k) { return imag ? (*imag)[k] : 0.0; };
    // a later entry for the same basis state replaces an earlier one
    std::map<int, double> amp2;
    for (size_t k = 0; k < indexes.size(); k++) amp2[indexes[k]] = coeffs[k] * coeffs[k] + im(k) * im(k);
    double norm2 = 0;
    for (const auto &e : amp2) norm2 += e.second;
    const double dev = std::fabs(std::sqrt(norm2) - 1.0);
    if (dev > 1e-14) log_msg("Note: the initial state's norm differs from 1 by %g", dev);
    start_.qubits = numberOfQubits;
    start_.bits = indexes;
    start_.amps = coeffs;
    start_.ampsIm = imag ? *imag : std::vector<double>();
    start_.complex = imag != nullptr;

    const int half = numberOfQubits / 2;
    const uint32_t downMask = (1u << half) - 1;
    Occupation o;
    o.fixedTotal = true;
    o.fixedSpins = numberOfQubits % 2 == 0;
    bool any = false;
    for (size_t k = 0; k < indexes.size() && o.fixedTotal; k++) {
        if (coeffs[k] == 0.0 && im(k) == 0.0) continue;
        const uint32_t b = (uint32_t)indexes[k];
        const int tot = __builtin_popcount(b), dn = __builtin_popcount(b & downMask),
                  up = __builtin_popcount(b >> half);
        if (!any) {
            o.total = tot;
            o.up = up;
            o.down = dn;
            any = true;
            continue;
        }
        if (tot != o.total) o.fixedTotal = false;
        else if (up != o.up || dn != o.down) o.fixedSpins = false;
    }
    if (!any) {
        log_msg("Note: the initial state has no nonzero component; the full space is used");
        o = Occupation{};
    } else if (!o.fixedTotal) {
        log_msg("Note: the initial state mixes particle numbers; the full space is used");
        o = Occupation{};
    } else if (!o.fixedSpins) {
        o.up = o.down = -1;
    }
    occ_ = o;
    return true;
}

bool Manager::storeHamiltonian(std::vector<int> &&i0, std::vector<int> &&j0, std::vector<double> &&v)
{
    bool ok = !frozen("Hamiltonian");
    if (i0.empty() || j0.empty() || v.empty()) {
        log_msg("Hamiltonian rejected: no entries");
        ok = false;
    } else if (i0.size() != j0.size() || i0.size() != v.size()) {
        log_msg("Hamiltonian rejected: %zu row, %zu column and %zu value entries", i0.size(), j0.size(), v.size());
        ok = false;
    }
    if (!ok) return false;
    trip_.i = std::move(i0);
    trip_.j = std::move(j0);
    trip_.v = std::move(v);
    return true;
}

bool Manager::storeIntegrals(int norb, const double *h, const double *g)
{
    if (frozen("Integrals")) return false;
    norb_ = norb;
    const size_t n2 = (size_t)norb * norb;
    h1_.assign(h, h + n2);
    g2_.assign(g, g + n2 * n2);
    return true;
}

bool Manager::storeRunPath(const std::string &runPath)
{
    const bool locked = frozen("Run path");
    if (runPath.empty()) log_msg("Run path rejected: empty");
    if (locked || runPath.empty()) return false;
    runPrefix_ = runPath;
    return true;
}

bool Manager::storeNuclearEnergy(double e)
{
    if (frozen("Nuclear energy")) return false;
    enuc_ = e;
    return true;
}

bool Manager::storeHamiltonianStorage(int mode)
{
    if (frozen("Hamiltonian storage")) return false;
    storage_ = mode;
    return true;
}

// Operator k takes angle scale_k * theta[param_k]; there are as many
// parameters as the largest index + 1, and no more parameters than
// operators (src/AnsatzManager.cpp:415-445).
bool Manager::storeParameterDependencies(const std::vector<std::pair<int, double>> &dep)
{
    const bool locked = frozen("Parameter map");
    int largest = -1;
    bool negative = false;
    for (const auto &d : dep) {
        largest = std::max(largest, d.first);
        negative = negative || d.first < 0;
    }
    if (dep.empty()) log_msg("Parameter map rejected: empty");
    if (negative) log_msg("Parameter map rejected: a parameter index below 1");
    if (locked || dep.empty() || negative) return false;
    paramIdx_.clear();
    paramScale_.clear();
    nparams_ = -1;
    if ((size_t)largest + 1 > dep.size()) {
        log_msg("Parameter map rejected: parameter %d for only %zu operators", largest + 1, dep.size());
        return false;
    }
    for (const auto &d : dep) {
        paramIdx_.push_back(d.first);
        paramScale_.push_back(d.second);
    }
    nparams_ = largest + 1;
    return true;
}

// ---------------------------------------------------------------- build

int Manager::chooseSector() const
{
    if (!occ_.fixedTotal) return kFull;
    if (!occ_.fixedSpins || start_.qubits % 2 != 0) return kNumber;
    // every operator must move as many up (down) electrons in as out
    // (src/AnsatzManager.cpp:202-237); an operator of another shape keeps
    // the S_z sector out and is rejected by the engine later
    const int q = start_.qubits;
    const uint32_t down = (1u << (q / 2)) - 1;
    for (const auto &e : ops_) {
        const bool single = e[0] >= 0 && e[1] >= 0 && e[2] < 0 && e[3] < 0;
        const bool dbl = e[0] >= 0 && e[1] >= 0 && e[2] >= 0 && e[3] >= 0;
        if ((!single && !dbl) || std::max(std::max(e[0], e[1]), std::max(e[2], e[3])) >= q) return kNumber;
        // created and annihilated orbitals
        const uint32_t in = dbl ? (1u << e[0]) | (1u << e[1]) : 1u << e[0];
        const uint32_t out = dbl ? (1u << e[2]) | (1u << e[3]) : 1u << e[1];
        if (__builtin_popcount(in & down) != __builtin_popcount(out & down) ||
            __builtin_popcount(in & ~down) != __builtin_popcount(out & ~down))
            return kNumber;
    }
    return ops_.empty() ? kNumber : kSz;
}

bool Manager::build()
{
    std::vector<const char *> missing;
    if (start_.qubits < 0 || start_.bits.empty()) missing.push_back("the initial state");
    if (ops_.empty()) missing.push_back("the operators");
    if (nparams_ < 0) missing.push_back("the parameter map (order file)");
    if (nparams_ >= 0 && !ops_.empty() && paramIdx_.size() != ops_.size())
        log_msg("Cannot build: %zu operators but %zu parameter map entries", ops_.size(), paramIdx_.size());
    if (!missing.empty()) {
        std::string m = missing[0];
        for (size_t k = 1; k < missing.size(); k++) m += std::string(", ") + missing[k];
        log_msg("Cannot build: missing %s", m.c_str());
    }
    if (!missing.empty() || paramIdx_.size() != ops_.size()) return false;

    // a run path stands in for a Hamiltonian that was not given directly
    if (trip_.i.empty() && h1_.empty() && !runPrefix_.empty()) {
        int kind = 0;
        std::vector<int> i0, j0;
        std::vector<double> v, h, g;
        std::string ferr;
        if (!files::read_run_hamiltonian(runPrefix_, start_.qubits, kind, i0, j0, v, h, g, ferr)) {
            log_msg("Cannot build: the run path's Hamiltonian: %s", ferr.c_str());
        } else if (kind == 1) {
            trip_.i = std::move(i0);
            trip_.j = std::move(j0);
            trip_.v = std::move(v);
        } else {
            norb_ = start_.qubits / 2;
            h1_ = std::move(h);
            g2_ = std::move(g);
        }
    }
    if (trip_.i.empty() && h1_.empty()) {
        log_msg("Cannot build: missing the Hamiltonian");
        return false;
    }

    auto eng = std::make_unique<Engine>();
    std::string err;
    if (!eng->init(start_.qubits, chooseSector(), occ_.total, occ_.up, occ_.down, err)) {
        log_msg("Device set-up failed: %s", err.c_str());
        return false;
    }
    const bool cplx = complexAmplitudes();
    if (cplx && !eng->set_complex(true, err)) {
        log_msg("Device set-up failed: %s", err.c_str());
        return false;
    }
    // operators and the parameter map (src/tupsquantities.cpp:643-667)
    std::vector<int8_t> flat;
    flat.reserve(4 * ops_.size());
    for (const auto &e : ops_) flat.insert(flat.end(), e.begin(), e.end());
    if (!eng->set_operators(flat, paramIdx_, paramScale_, nparams_, err)) {
        log_msg("Operator set-up failed: %s", err.c_str());
        return false;
    }
    // the start state over the sector (components outside it are dropped,
    // as compressVector does: src/sparsematrix.h:83-105)
    std::vector<double> psi0(eng->dim(), 0.0), psi0i(cplx ? eng->dim() : 0, 0.0);
    for (size_t k = 0; k < start_.bits.size(); k++) {
        const long long r = eng->rank((uint32_t)start_.bits[k]);
        if (r < 0) continue;
        psi0[r] = start_.amps[k];
        if (cplx && !start_.ampsIm.empty()) psi0i[r] = start_.ampsIm[k];
    }
    if (!(cplx ? eng->set_start_complex(psi0, psi0i, err) : eng->set_start(psi0, err))) {
        log_msg("Start state upload failed: %s", err.c_str());
        return false;
    }
    if (!trip_.i.empty()) {
        // rows of y = M^T x over the sector: output row j holds (i, v) for
        // every stored (i, j, v), repeated (i, j) summed, entries with an
        // end outside the sector dropped (src/hamiltonianmatrix.cpp:628-659)
        const long long N = eng->dim();
        std::vector<std::pair<long long, std::pair<int32_t, double>>> ent;
        ent.reserve(trip_.i.size());
        for (size_t k = 0; k < trip_.i.size(); k++) {
            if (trip_.i[k] < 0 || trip_.j[k] < 0) continue;
            const long long ri = eng->rank((uint32_t)trip_.i[k]), rj = eng->rank((uint32_t)trip_.j[k]);
            if (ri >= 0 && rj >= 0) ent.push_back({rj, {(int32_t)ri, trip_.v[k]}});
        }
        std::stable_sort(ent.begin(), ent.end(), [](const auto &x, const auto &y) {
            return x.first != y.first ? x.first < y.first : x.second.first < y.second.first;
        });
        std::vector<int64_t> rowptr(N + 1, 0);
        std::vector<int32_t> col;
        std::vector<double> val;
        for (size_t k = 0; k < ent.size(); k++) {
            const bool repeat = k > 0 && ent[k].first == ent[k - 1].first && ent[k].second.first == ent[k - 1].second.first;
            if (repeat) {
                val.back() += ent[k].second.second;
                continue;
            }
            col.push_back(ent[k].second.first);
            val.push_back(ent[k].second.second);
            rowptr[ent[k].first + 1]++;
        }
        for (long long r = 0; r < N; r++) rowptr[r + 1] += rowptr[r];
        if (!eng->set_ham_rows(rowptr, col, val, err)) {
            log_msg("Hamiltonian upload failed: %s", err.c_str());
            return false;
        }
        trip_ = Triplets{};
    } else {
        if (2 * norb_ != start_.qubits) {
            log_msg("Cannot build: integrals for %d orbitals against %d qubits", norb_, start_.qubits);
            return false;
        }
        if (!eng->set_ham_integrals(norb_, h1_, g2_, storage_, err)) {
            log_msg("Integral Hamiltonian set-up failed: %s", err.c_str());
            return false;
        }
    }
    // the state at zero angles, as the reference evolves on construction
    std::vector<double> zero(ops_.size(), 0.0);
    if (!eng->evolve(zero.data(), err)) {
        log_msg("Evolution failed: %s", err.c_str());
        return false;
    }
    cur_ = zero;
    eng_ = std::move(eng);
    return true;
}

bool Manager::ensureBuilt()
{
    if (eng_) return true;
    if (build()) return true;
    log_msg("Evaluation refused: the set-up is incomplete or could not be built");
    return false;
}

// ---------------------------------------------------------------- angles

bool Manager::toOperatorAngles(const double *in, size_t n, double *out) const
{
    const size_t M = ops_.size();
    if (n == M) {
        std::copy(in, in + M, out);
        return true;
    }
    if ((int)n == nparams_) {
        for (size_t k = 0; k < M; k++) out[k] = paramScale_[k] * in[paramIdx_[k]];
        return true;
    }
    log_msg("Angles rejected: %zu given, %zu per operator or %d per parameter expected", n, M, nparams_);
    return false;
}

bool Manager::toOperatorAngles(const std::vector<double> &in, std::vector<double> &out) const
{
    out.resize(ops_.size());
    return toOperatorAngles(in.data(), in.size(), out.data());
}

// Per-operator or per-parameter angles (src/AnsatzManager.cpp:467-506); psi
// is evolved again only when the per-operator angles change.
bool Manager::setAngles(std::vector<double> angles)
{
    if (!ensureBuilt()) return false;
    std::vector<double> x;
    if (!toOperatorAngles(angles, x)) return false;
    if (x == cur_) return true;
    std::string err;
    if (!eng_->evolve(x.data(), err)) {
        log_msg("Evolution failed: %s", err.c_str());
        cur_.assign(ops_.size(), std::nan(""));
        return false;
    }
    cur_ = x;
    return true;
}

void Manager::compressGradient(const double *deriv, double *grad) const
{
    std::fill(grad, grad + nparams_, 0.0);
    for (size_t k = 0; k < paramIdx_.size(); k++) grad[paramIdx_[k]] += paramScale_[k] * deriv[k];
}

bool Manager::getExpectationValue(const std::vector<double> &angles, double &E)
{
    if (!setAngles(angles)) return false;
    std::string err;
    if (!eng_->energy(&E, err)) {
        log_msg("Energy failed: %s", err.c_str());
        return false;
    }
    E += enuc_;
    return true;
}

bool Manager::realOnly(const char *what) const
{
    if (!eng_ || !eng_->is_complex()) return true;
    log_msg("%s is not available for complex amplitudes", what);
    return false;
}

bool Manager::complexAmplitudes() const
{
    if (start_.complex) return true;
    for (const auto &e : ops_) {
        const bool single = e[2] < 0 && e[3] < 0;
        const uint32_t cre = single ? 1u << e[0] : (1u << e[0]) | (1u << e[1]);
        const uint32_t ann = single ? 1u << e[1] : (1u << e[2]) | (1u << e[3]);
        if (e[0] >= 0 && e[1] >= 0 && cre == ann) return true;  // a number operator
    }
    return false;
}

bool Manager::getFinalStateComplex(const std::vector<double> &angles, std::vector<double> &re,
                                   std::vector<double> &im)
{
    if (!setAngles(angles)) return false;
    std::string err;
    std::vector<double> cr(eng_->dim()), ci(eng_->dim());
    if (!eng_->download_state_complex(cr.data(), ci.data(), err)) {
        log_msg("State download failed: %s", err.c_str());
        return false;
    }
    re.assign((size_t)1 << start_.qubits, 0.0);
    im.assign(re.size(), 0.0);
    const auto &st = eng_->states();
    for (size_t i = 0; i < cr.size(); i++) {
        re[st[i]] = cr[i];
        im[st[i]] = ci[i];
    }
    return true;
}

bool Manager::getFinalState(const std::vector<double> &angles, std::vector<double> &full)
{
    if (!setAngles(angles)) return false;
    if (eng_->is_complex()) log_msg("Note: complex amplitudes; the final state returned is their real part");
    std::string err;
    std::vector<double> comp(eng_->dim());
    if (!eng_->download_state(comp.data(), err)) {
        log_msg("State download failed: %s", err.c_str());
        return false;
    }
    full.assign((size_t)1 << start_.qubits, 0.0);
    const auto &st = eng_->states();
    for (size_t i = 0; i < comp.size(); i++) full[st[i]] = comp[i];
    return true;
}

// Decompresses `angles` like setAngles; if they differ from the cached
// angles, evolves and differentiates in one fused device sequence,
// otherwise differentiates the cached psi (evolveDerivative,
// src/fusedevolve.cpp:2263-2346).
bool Manager::gradientAt(const std::vector<double> &angles, double *E, std::vector<double> &deriv)
{
    if (!ensureBuilt()) return false;
    const size_t M = ops_.size();
    std::vector<double> x;
    if (!toOperatorAngles(angles, x)) return false;
    std::string err;
    deriv.resize(M);
    bool ok;
    if (x != cur_) {
        ok = eng_->energy_gradient(x.data(), E, deriv.data(), err);
        cur_ = ok ? x : std::vector<double>(M, std::nan(""));
    } else {
        ok = eng_->gradient(x.data(), E, deriv.data(), err);
    }
    if (!ok) log_msg("Gradient failed: %s", err.c_str());
    return ok;
}

bool Manager::getGradientComp(const std::vector<double> &angles, std::vector<double> &grad)
{
    std::vector<double> deriv;
    if (!gradientAt(angles, nullptr, deriv)) return false;
    grad.resize(nparams_);
    compressGradient(deriv.data(), grad.data());
    return true;
}

bool Manager::getGradient(const std::vector<double> &angles, std::vector<double> &grad)
{
    return gradientAt(angles, nullptr, grad);
}

bool Manager::getOverlapGradient(const std::vector<double> &angles, const std::vector<double> &projFull,
                                 double &overlap, std::vector<double> &grad)
{
    if (!ensureBuilt()) return false;
    const size_t M = ops_.size();
    std::vector<double> x;
    if (!toOperatorAngles(angles, x)) return false;
    if (projFull.size() != ((size_t)1 << start_.qubits)) {
        log_msg("Projection vector has %zu entries, expected 2^%d", projFull.size(), start_.qubits);
        return false;
    }
    // the sector's amplitudes of proj (entries outside the sector meet zeros of psi)
    const auto &st = eng_->states();
    std::vector<double> comp(st.size());
    for (size_t i = 0; i < st.size(); i++) comp[i] = projFull[st[i]];
    std::vector<double> deriv(M);
    std::string err;
    if (!eng_->overlap_gradient(x.data(), comp.data(), &overlap, deriv.data(), err)) {
        log_msg("Overlap gradient failed: %s", err.c_str());
        cur_.assign(M, std::nan(""));
        return false;
    }
    cur_ = x;  // psi is now U(x) psi_0
    grad.resize(nparams_);
    compressGradient(deriv.data(), grad.data());
    return true;
}

bool Manager::getEnergyGradient(const std::vector<double> &angles, double &E, std::vector<double> &grad)
{
    std::vector<double> deriv;
    if (!gradientAt(angles, &E, deriv)) return false;
    E += enuc_;
    grad.resize(nparams_);
    compressGradient(deriv.data(), grad.data());
    return true;
}

int Manager::getEnergyGradientNoAlloc(const double *angles, int nangles, double &E, double *grad)
{
    if (!eng_) {
        log_msg("No-allocation evaluation refused: the context is not built yet (an ordinary evaluation builds it)");
        return 1;
    }
    if (eng_->noalloc_scratch_need() > Engine::noalloc_scratch_doubles()) {
        log_msg("No-allocation evaluation refused: the set-up needs %zu doubles of scratch, at most %zu are there",
                eng_->noalloc_scratch_need(), Engine::noalloc_scratch_doubles());
        return 2;
    }
    // per-operator angles into the reused member (its capacity stays after the first call)
    nostage_.resize(ops_.size());
    if (!toOperatorAngles(angles, (size_t)nangles, nostage_.data())) return 3;
    noderiv_.resize(ops_.size());
    std::string err;
    if (!eng_->energy_gradient_noalloc(nostage_.data(), &E, noderiv_.data(), err)) {
        log_msg("No-allocation evaluation failed: %s", err.c_str());
        cur_.assign(ops_.size(), std::nan(""));
        return 3;
    }
    cur_ = nostage_;
    E += enuc_;
    compressGradient(noderiv_.data(), grad);
    return 0;
}

bool Manager::getHessianComp(const std::vector<double> &angles, std::vector<double> &hess)
{
    if (!setAngles(angles)) return false;
    std::string err;
    const int P = nparams_;
    hess.assign((size_t)P * P, 0.0);
    std::vector<double> g(P);
    double E;
    if (!eng_->hessian(cur_.data(), hess.data(), g.data(), &E, err)) {
        log_msg("Hessian failed: %s", err.c_str());
        return false;
    }
    return true;
}

bool Manager::getHessianProj(const std::vector<double> &angles, const std::vector<double> &projFull,
                             std::vector<double> &hess, std::vector<double> &grad, double &overlap)
{
    if (!ensureBuilt()) return false;
    std::vector<double> x;
    if (!toOperatorAngles(angles, x)) return false;
    if (projFull.size() != ((size_t)1 << start_.qubits)) {
        log_msg("Projection vector has %zu entries, expected 2^%d", projFull.size(), start_.qubits);
        return false;
    }
    const auto &st = eng_->states();
    std::vector<double> comp(st.size());
    for (size_t i = 0; i < st.size(); i++) comp[i] = projFull[st[i]];
    const int P = nparams_;
    hess.assign((size_t)P * P, 0.0);
    grad.assign(P, 0.0);
    std::string err;
    // the projected Hessian runs its own forward sweep; psi ends as U(x) psi_0
    if (!eng_->hessian_proj(x.data(), comp.data(), hess.data(), grad.data(), &overlap, err)) {
        log_msg("Projected Hessian failed: %s", err.c_str());
        cur_.assign(ops_.size(), std::nan(""));
        return false;
    }
    cur_ = x;
    return true;
}

bool Manager::decompressBatch(int nvec, int nangles, const double *angles, std::vector<double> &out)
{
    const size_t M = ops_.size();
    out.resize((size_t)nvec * M);
    for (int v = 0; v < nvec; v++)
        if (!toOperatorAngles(angles + (size_t)v * nangles, (size_t)nangles, out.data() + (size_t)v * M)) return false;
    return true;
}

// src/AnsatzManager.cpp:709-754
bool Manager::getExpectationValues(int nvec, int nangles, const double *angles, std::vector<double> &E)
{
    if (!ensureBuilt()) return false;
    std::vector<double> dec;
    if (!decompressBatch(nvec, nangles, angles, dec)) return false;
    E.resize(nvec);
    std::string err;
    if (!eng_->batch(nvec, dec.data(), E.data(), nullptr, err)) {
        log_msg("Batched energies failed: %s", err.c_str());
        return false;
    }
    for (auto &x : E) x += enuc_;
    return true;
}

bool Manager::getEnergyGradients(int nvec, int nangles, const double *angles, std::vector<double> &E,
                                 std::vector<double> &grads)
{
    if (!ensureBuilt()) return false;
    std::vector<double> dec;
    if (!decompressBatch(nvec, nangles, angles, dec)) return false;
    const size_t M = ops_.size();
    E.resize(nvec);
    std::vector<double> deriv((size_t)nvec * M);
    std::string err;
    if (!eng_->batch(nvec, dec.data(), E.data(), deriv.data(), err)) {
        log_msg("Batched gradients failed: %s", err.c_str());
        return false;
    }
    grads.resize((size_t)nvec * nparams_);
    for (int v = 0; v < nvec; v++) {
        E[v] += enuc_;
        compressGradient(deriv.data() + (size_t)v * M, grads.data() + (size_t)v * nparams_);
    }
    return true;
}

bool Manager::uploadAngleBatch(int nvec, int nangles, const double *angles)
{
    if (!ensureBuilt()) return false;
    std::vector<double> dec;
    if (!decompressBatch(nvec, nangles, angles, dec)) return false;
    std::string err;
    if (!eng_->batch_upload(nvec, dec.data(), err)) {
        log_msg("Batch upload failed: %s", err.c_str());
        return false;
    }
    batchNangles_ = nangles;
    return true;
}

bool Manager::runAngleBatch(bool withGradient)
{
    if (!eng_ || eng_->batch_size() < 1) {
        log_msg("No angle batch uploaded");
        return false;
    }
    std::string err;
    if (!eng_->batch_run(withGradient, err)) {
        log_msg("Batch run failed: %s", err.c_str());
        return false;
    }
    return true;
}

bool Manager::downloadAngleBatch(double *energies, double *gradients)
{
    if (!eng_ || eng_->batch_size() < 1) {
        log_msg("No angle batch uploaded");
        return false;
    }
    const int nvec = eng_->batch_size();
    const size_t M = ops_.size();
    std::vector<double> deriv(gradients ? (size_t)nvec * M : 0);
    std::string err;
    if (!eng_->batch_download(energies, gradients ? deriv.data() : nullptr, err)) {
        log_msg("Batch download failed: %s", err.c_str());
        return false;
    }
    for (int v = 0; v < nvec; v++) {
        energies[v] += enuc_;
        if (gradients)
            compressGradient(deriv.data() + (size_t)v * M, gradients + (size_t)v * nparams_);
    }
    return true;
}

bool Manager::setKernelTiming(bool on)
{
    if (!ensureBuilt()) return false;
    eng_->set_timing(on);
    return true;
}

bool Manager::getKernelTiming(int cls, double &ms, long long &launches)
{
    if (!ensureBuilt()) return false;
    std::string err;
    if (!eng_->timing(cls, &ms, &launches, err)) {
        log_msg("Timing query failed: %s", err.c_str());
        return false;
    }
    return true;
}

bool Manager::getDimensions(long long &dim, int &nops, int &nparams)
{
    if (!ensureBuilt()) return false;
    dim = eng_->dim();
    nops = (int)ops_.size();
    nparams = nparams_;
    return true;
}

} // namespace ae

namespace ae {

bool Manager::optimise(std::vector<double> &angles, double &E, std::vector<double> &trace, double stopEnergy,
                       bool useStop)
{
    if (!ensureBuilt() || !realOnly("The Newton optimiser")) return false;
    const size_t M = ops_.size();
    const int P = nparams_;
    const bool perParam = (int)angles.size() == P && angles.size() != M;
    std::vector<double> x;
    if (!toOperatorAngles(angles, x)) return false;
    Engine *eng = eng_.get();
    std::string err;
    EvalFns f;
    f.hessian = [&](const std::vector<double> &a, std::vector<double> &H, std::vector<double> &g, double &En) {
        if (!eng->hessian(a.data(), H.data(), g.data(), &En, err)) return false;
        En += enuc_;
        return true;
    };
    f.egrad = [&](const std::vector<double> &a, double &En, std::vector<double> &g) {
        std::vector<double> deriv(M);
        if (!eng->evolve(a.data(), err) || !eng->gradient(a.data(), &En, deriv.data(), err)) return false;
        En += enuc_;
        compressGradient(deriv.data(), g.data());
        return true;
    };
    NewtonDriver nd(paramIdx_, paramScale_, P, f);
    if (useStop) nd.set_stop_energy(stopEnergy);
    const bool ok = nd.optimise(x, E);
    // psi now belongs to the last evaluated angles; force a re-evolve next time
    cur_.assign(M, std::nan(""));
    if (!ok) {
        log_msg("Optimisation failed: %s", err.c_str());
        return false;
    }
    trace.clear();
    for (const auto &tp : nd.trace) {
        trace.push_back(tp.seconds);
        trace.push_back(tp.energy);
        trace.push_back(tp.gradNorm);
        trace.push_back((double)tp.hessianEvals);
        trace.push_back((double)tp.energyEvals);
    }
    if (perParam) {
        // normCompress: first operator of each parameter, divided by its scale
        // (src/tupsquantities.cpp:655-666)
        std::vector<bool> seen(P, false);
        for (size_t k = 0; k < M; k++) {
            const int p = paramIdx_[k];
            if (!seen[p]) {
                angles[p] = x[k] / paramScale_[k];
                seen[p] = true;
            }
        }
    } else {
        angles = x;
    }
    return true;
}

bool Manager::iterativeOptimise(std::vector<double> &angles, double &E, std::vector<double> &trace, bool ediis)
{
    if (!ensureBuilt() || !realOnly("The iterative optimiser")) return false;
    const size_t M = ops_.size();
    const int P = nparams_;
    const bool perParam = (int)angles.size() == P && angles.size() != M;
    std::vector<double> x;
    if (!toOperatorAngles(angles, x)) return false;
    Engine *eng = eng_.get();
    const auto &st = eng->states();
    std::vector<double> proj(st.size());  // H psi of the last outer iterate, sector order
    std::string err;
    ProjFns f;
    f.setProjection = [&](const std::vector<double> &a, std::vector<double> &psi) {
        psi.resize(st.size());
        return eng->evolve(a.data(), err) && eng->download_state(psi.data(), err) &&
               eng->download_ham_state(proj.data(), err);
    };
    f.hessianProj = [&](const std::vector<double> &a, std::vector<double> &H, std::vector<double> &g, double &v) {
        H.assign((size_t)P * P, 0.0);
        g.assign(P, 0.0);
        return eng->hessian_proj(a.data(), proj.data(), H.data(), g.data(), &v, err);
    };
    f.gradProj = [&](const std::vector<double> &a, double &v, std::vector<double> &g) {
        std::vector<double> deriv(M);
        if (!eng->overlap_gradient(a.data(), proj.data(), &v, deriv.data(), err)) return false;
        g.resize(P);
        compressGradient(deriv.data(), g.data());
        return true;
    };
    ProjectedDriver pd(paramIdx_, paramScale_, P, f);
    pd.useEdiis = ediis;
    const bool ok = pd.optimise(x, E);
    cur_.assign(M, std::nan(""));  // psi belongs to the last evaluated angles
    if (!ok) {
        log_msg("Iterative optimisation failed: %s", err.c_str());
        return false;
    }
    E += enuc_;
    trace.clear();
    for (const auto &tp : pd.trace) {
        trace.push_back(tp.seconds);
        trace.push_back(tp.energy + enuc_);
        trace.push_back(tp.gradNorm);
        trace.push_back((double)tp.hessianEvals);
        trace.push_back((double)tp.energyEvals);
    }
    if (perParam) {
        std::vector<bool> seen(P, false);
        for (size_t k = 0; k < M; k++)
            if (!seen[paramIdx_[k]]) {
                angles[paramIdx_[k]] = x[k] / paramScale_[k];
                seen[paramIdx_[k]] = true;
            }
    } else {
        angles = x;
    }
    return true;
}

bool Manager::subspaceDiagonalise(int nsets, int nangles, const double *angles, std::vector<double> &evals,
                                  std::vector<double> &H, std::vector<double> &S, int &nullRank)
{
    if (!ensureBuilt() || !realOnly("Subspace diagonalisation")) return false;
    std::vector<double> dec;
    if (!decompressBatch(nsets, nangles, angles, dec)) return false;
    H.assign((size_t)nsets * nsets, 0.0);
    S.assign((size_t)nsets * nsets, 0.0);
    std::string err;
    if (!eng_->subspace(nsets, dec.data(), H.data(), S.data(), err)) {
        log_msg("Subspace Gram matrices failed: %s", err.c_str());
        return false;
    }
    // exact symmetry: <i|H|j> and <j|H|i> are summed in different orders
    for (int i = 0; i < nsets; i++)
        for (int j = 0; j < i; j++) {
            const double h = 0.5 * (H[(size_t)i * nsets + j] + H[(size_t)j * nsets + i]);
            const double o = 0.5 * (S[(size_t)i * nsets + j] + S[(size_t)j * nsets + i]);
            H[(size_t)i * nsets + j] = H[(size_t)j * nsets + i] = h;
            S[(size_t)i * nsets + j] = S[(size_t)j * nsets + i] = o;
        }
    subspace_eigen(nsets, H, S, evals, nullRank);
    return true;
}

// Per angle vector (src/tupsquantities.cpp:185-568): energies, counts of
// negative / near-zero / positive eigenvalues and diagonal entries of the
// Hessian (|e| <= 1e-10 is near zero), of positive and near-zero
// eigenvalues of the metric g_mu,nu = <dpsi/dtheta_mu|dpsi/dtheta_nu> and of
// its near-zero diagonal entries, the gradient norm, and the overlaps with
// the lowest eigenvectors of H.
bool Manager::properties(int nsets, int nangles, const double *angles, int noverlaps, bool noHess,
                         std::vector<double> &rows, std::vector<double> &hessians, std::vector<double> &metrics,
                         std::vector<double> &lowest)
{
    if (!ensureBuilt() || !realOnly("The properties table")) return false;
    const int P = nparams_;
    const long long N = eng_->dim();
    const int cols = kPropColumns + 2 * noverlaps;
    std::string err;
    std::vector<double> vecs;
    lowest.clear();
    if (noverlaps > 0) {
        std::vector<double> start(N, 0.0);
        for (size_t k = 0; k < start_.bits.size(); k++) {
            const long long r = eng_->rank((uint32_t)start_.bits[k]);
            if (r >= 0) start[r] = start_.amps[k];
        }
        auto applyH = [&](const double *x, double *y) { return eng_->apply_ham(x, y, err); };
        if (!lowest_eigenpairs(noverlaps, N, applyH, start, lowest, vecs, 400, 1e-10)) {
            log_msg("Lowest eigenvectors did not converge: %s", err.c_str());
            return false;
        }
    }
    rows.assign((size_t)nsets * cols, 0.0);
    hessians.assign(noHess ? 0 : (size_t)nsets * P * P, 0.0);
    metrics.assign(noHess ? 0 : (size_t)nsets * P * P, 0.0);
    const double zero = 1e-10;
    std::vector<double> x, psi(N), g(P), Hm((size_t)P * P), Gm((size_t)P * P), w, V;
    for (int s = 0; s < nsets; s++) {
        x.resize(ops_.size());
        if (!toOperatorAngles(angles + (size_t)s * nangles, (size_t)nangles, x.data())) return false;
        double *row = rows.data() + (size_t)s * cols;
        double E = 0;
        if (!noHess) {
            if (!eng_->hessian(x.data(), Hm.data(), g.data(), &E, err, Gm.data())) {
                log_msg("Hessian failed: %s", err.c_str());
                cur_.assign(ops_.size(), std::nan(""));
                return false;
            }
            std::copy(Hm.begin(), Hm.end(), hessians.begin() + (size_t)s * P * P);
            std::copy(Gm.begin(), Gm.end(), metrics.begin() + (size_t)s * P * P);
            symmetric_eigen(P, Hm, w, V);
            for (double e : w) row[e < -zero ? 2 : e > zero ? 4 : 3] += 1;
            for (int i = 0; i < P; i++) {
                const double d = Hm[(size_t)i * P + i];
                row[d < -zero ? 5 : d > zero ? 7 : 6] += 1;
            }
            symmetric_eigen(P, Gm, w, V);
            for (double e : w) row[e > zero ? 8 : 9] += 1;  // a metric has no negative eigenvalues
            for (int i = 0; i < P; i++) row[10] += std::fabs(Gm[(size_t)i * P + i]) <= zero ? 1 : 0;
        } else {
            std::vector<double> deriv(ops_.size());
            if (!eng_->energy_gradient(x.data(), &E, deriv.data(), err)) {
                log_msg("Gradient failed: %s", err.c_str());
                cur_.assign(ops_.size(), std::nan(""));
                return false;
            }
            compressGradient(deriv.data(), g.data());
        }
        cur_.assign(ops_.size(), std::nan(""));
        row[0] = E;
        row[1] = E + enuc_;
        double gn = 0;
        for (double v : g) gn += v * v;
        row[11] = std::sqrt(gn);
        if (noverlaps > 0) {
            if (!eng_->evolve(x.data(), err) || !eng_->download_state(psi.data(), err)) {
                log_msg("State failed: %s", err.c_str());
                return false;
            }
            cur_ = x;
            for (int t = 0; t < noverlaps; t++) {
                double o = 0;
                for (long long i = 0; i < N; i++) o += psi[i] * vecs[(size_t)t * N + i];
                row[kPropColumns + 2 * t] = o;
                row[kPropColumns + 2 * t + 1] = std::fabs(o);
            }
        }
    }
    return true;
}

} // namespace ae
