// Host-side mirror of stateAnsatzManager (src/AnsatzManager.cpp).
#include "manager.hpp"
#include "runfiles.hpp"

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

bool Manager::storeOperators(const std::vector<std::array<int8_t, 4>> &excs)
{
    bool success = true;
    if (m_isConstructed) {
        success = false;
        log_msg("Already constructed, setup cant be modified now");
    }
    if (excs.empty()) {
        success = false;
        log_msg("Excitations given has zero size");
    }
    if (success) m_excitations = excs;
    return success;
}

// src/AnsatzManager.cpp:271-374: particle number and S_z of the non-zero
// components decide which sector the state lives in.
bool Manager::storeInitial(int numberOfQubits, const std::vector<int> &indexes, const std::vector<double> &coeffs)
{
    bool success = true;
    if (m_isConstructed) {
        success = false;
        log_msg("Already constructed, setup cant be modified now");
    }
    if (numberOfQubits < 1) {
        success = false;
        log_msg("Number of qubits < 1: %d", numberOfQubits);
    }
    if (numberOfQubits > 30) {
        success = false;
        log_msg("Number of qubits > 30 is not supported: %d", numberOfQubits);
    }
    if (indexes.empty() || coeffs.empty()) {
        success = false;
        log_msg("Initial state has no elements");
    }
    if (coeffs.size() != indexes.size()) {
        success = false;
        log_msg("Initial state indexes and coeffs are a different size");
    }
    for (int idx : indexes)
        if (success && (idx < 0 || (numberOfQubits <= 30 && (long long)idx >= (1ll << numberOfQubits)))) {
            success = false;
            log_msg("Initial state index out of range: %d", idx + 1);
        }
    if (!success) return false;
    // the dense start vector of the reference: later duplicates overwrite
    std::map<int, double> dense;
    for (size_t i = 0; i < indexes.size(); i++) dense[indexes[i]] = coeffs[i];
    double mag = 0;
    for (auto &kv : dense) mag += kv.second * kv.second;
    mag = std::sqrt(mag);
    if (std::abs(mag - 1) > 1e-14) log_msg("Warning: Initial vector is not normalised, the error is %g", std::abs(mag - 1));
    m_startIdx = indexes;
    m_startCoeff = coeffs;
    m_numberOfQubits = numberOfQubits;

    int numberOfParticles = -1, spinUp = -1, spinDown = -1;
    bool allSame = true, SZSym = (numberOfQubits % 2 == 0);
    const uint32_t low = (numberOfQubits / 2 >= 32) ? 0xffffffffu : ((1u << (numberOfQubits / 2)) - 1);
    for (size_t i = 0; i < indexes.size(); i++) {
        if (coeffs[i] == 0.) continue;
        const uint32_t s = (uint32_t)indexes[i];
        const int c = __builtin_popcount(s);
        if (numberOfParticles == -1) numberOfParticles = c;
        if (c != numberOfParticles) {
            allSame = false;
            break;
        }
        const int cu = __builtin_popcount(s >> (numberOfQubits / 2));
        const int cd = __builtin_popcount(s & low);
        if (spinUp == -1) spinUp = cu;
        if (spinDown == -1) spinDown = cd;
        if (cu != spinUp || cd != spinDown) SZSym = false;
    }
    if (numberOfParticles == -1) log_msg("Warning: Could not determine particle number");
    else if (!allSame) log_msg("Warning: not all the same particle number");
    m_numberOfParticles = -1;
    m_spinUp = m_spinDown = -1;
    m_InitialSZSym = false;
    m_particleSym = false;
    if (allSame && numberOfParticles != -1) {
        m_numberOfParticles = numberOfParticles;
        m_particleSym = true;
        if (SZSym && spinUp != -1 && spinDown != -1) {
            m_InitialSZSym = true;
            m_spinUp = spinUp;
            m_spinDown = spinDown;
        }
    }
    return true;
}

bool Manager::storeHamiltonian(std::vector<int> &&i0, std::vector<int> &&j0, std::vector<double> &&v)
{
    bool success = true;
    if (m_isConstructed) {
        success = false;
        log_msg("Already constructed, setup cant be modified now");
    }
    if (i0.empty() || j0.empty() || v.empty()) {
        success = false;
        log_msg("Hamiltonian Matrix has no elements");
    }
    if (i0.size() != j0.size() || i0.size() != v.size()) {
        success = false;
        log_msg("Hamiltonian Matrix indexes and coeffs are a different size");
    }
    if (!success) return false;
    m_iIndexes = std::move(i0);
    m_jIndexes = std::move(j0);
    m_coeffs = std::move(v);
    return true;
}

bool Manager::storeIntegrals(int norb, const double *h, const double *g)
{
    if (m_isConstructed) {
        log_msg("Already constructed, setup cant be modified now");
        return false;
    }
    m_norb = norb;
    const size_t n2 = (size_t)norb * norb;
    m_h.assign(h, h + n2);
    m_g.assign(g, g + n2 * n2);
    return true;
}

bool Manager::storeRunPath(const std::string &runPath)
{
    bool success = true;
    if (m_isConstructed) {
        success = false;
        log_msg("Already constructed, setup cant be modified now");
    }
    if (runPath.empty()) {
        success = false;
        log_msg("Run path has zero size");
    }
    if (success) m_runPath = runPath;
    return success;
}

bool Manager::storeNuclearEnergy(double e)
{
    if (m_isConstructed) {
        log_msg("Already constructed, setup cant be modified now");
        return false;
    }
    m_nuclearEnergy = e;
    return true;
}

bool Manager::storeHamiltonianStorage(int mode)
{
    if (m_isConstructed) {
        log_msg("Already constructed, setup cant be modified now");
        return false;
    }
    m_hamStorage = mode;
    return true;
}

// src/AnsatzManager.cpp:415-445
bool Manager::storeParameterDependencies(const std::vector<std::pair<int, double>> &dep)
{
    bool success = true;
    if (m_isConstructed) {
        success = false;
        log_msg("Already constructed, setup cant be modified now");
    }
    if (dep.empty()) {
        success = false;
        log_msg("Parameter Dependencies, (order file) has zero size");
    }
    for (auto &d : dep)
        if (d.first < 0) {
            success = false;
            log_msg("Parameter index < 1 in the order file");
            break;
        }
    if (!success) return false;
    m_parameterDependency = dep;
    m_numberOfUniqueParameters = -1;
    for (auto &d : dep) m_numberOfUniqueParameters = std::max(d.first + 1, m_numberOfUniqueParameters);
    if ((size_t)m_numberOfUniqueParameters > dep.size()) {
        m_numberOfUniqueParameters = -1;
        m_parameterDependency.clear();
        return false;
    }
    return true;
}

// src/AnsatzManager.cpp:202-237
void Manager::setOperatorSymmetry()
{
    m_OperatorSZSym = true;
    const int nq = m_numberOfQubits;
    const uint32_t down = (1u << (nq / 2)) - 1;
    const uint32_t up = ((nq >= 32) ? 0xffffffffu : ((1u << nq) - 1)) ^ down;
    for (const auto &e : m_excitations) {
        uint32_t cr = 0, de = 0;
        if (e[0] >= nq || e[1] >= nq || e[2] >= nq || e[3] >= nq) {
            m_OperatorSZSym = false;  // out of range: rejected later by the engine
            continue;
        }
        if (e[0] >= 0 && e[1] >= 0 && e[2] < 0 && e[3] < 0) {
            cr = 1u << e[0];
            de = 1u << e[1];
        } else if (e[0] >= 0 && e[1] >= 0 && e[2] >= 0 && e[3] >= 0) {
            cr = (1u << e[0]) | (1u << e[1]);
            de = (1u << e[2]) | (1u << e[3]);
        } else {
            m_OperatorSZSym = false;  // rejected later by the engine
            continue;
        }
        m_OperatorSZSym = m_OperatorSZSym && __builtin_popcount(cr & up) == __builtin_popcount(de & up) &&
                          __builtin_popcount(cr & down) == __builtin_popcount(de & down);
    }
    if (m_excitations.empty()) m_OperatorSZSym = false;
}

bool Manager::construct()
{
    if (m_isConstructed) return true;
    bool success = true;
    if (m_numberOfQubits == -1) {
        success = false;
        log_msg("Number of Qubits is not set. Make sure the initial state has been provided and has elements");
    }
    if (m_startIdx.empty()) {
        success = false;
        log_msg("Make sure the initial state has been provided");
    }
    if (m_excitations.empty()) {
        success = false;
        log_msg("No excitations have been set");
    }
    if (m_parameterDependency.empty() || m_numberOfUniqueParameters == -1) {
        success = false;
        log_msg("Parameter Dependencies, (order file) has zero size");
    }
    if (m_parameterDependency.size() != m_excitations.size()) {
        success = false;
        log_msg("More operators than parameters dependencies given");
    }
    if (success && m_iIndexes.empty() && m_h.empty() && !m_runPath.empty()) {
        // the Hamiltonian from the run path's files
        int kind = 0;
        std::vector<int> i0, j0;
        std::vector<double> v, h, g;
        std::string ferr;
        if (!files::read_run_hamiltonian(m_runPath, m_numberOfQubits, kind, i0, j0, v, h, g, ferr)) {
            log_msg("Could not construct Hamiltonian: %s", ferr.c_str());
        } else if (kind == 1) {
            m_iIndexes = std::move(i0);
            m_jIndexes = std::move(j0);
            m_coeffs = std::move(v);
        } else {
            m_norb = m_numberOfQubits / 2;
            m_h = std::move(h);
            m_g = std::move(g);
        }
    }
    if (m_iIndexes.empty() && m_h.empty()) {
        success = false;
        log_msg("No Hamiltonian has been set");
    }
    if (!success) return false;
    setOperatorSymmetry();
    int mode = kFull;
    if (m_particleSym) mode = (m_InitialSZSym && m_OperatorSZSym) ? kSz : kNumber;
    auto eng = std::make_unique<Engine>();
    std::string err;
    if (!eng->init(m_numberOfQubits, mode, m_numberOfParticles, m_spinUp, m_spinDown, err)) {
        log_msg("Device set-up failed: %s", err.c_str());
        return false;
    }
    // operators and the parameter map (src/tupsquantities.cpp:643-667)
    std::vector<int8_t> ops4;
    std::vector<int> order;
    std::vector<double> scale;
    for (size_t k = 0; k < m_excitations.size(); k++) {
        for (int j = 0; j < 4; j++) ops4.push_back(m_excitations[k][j]);
        order.push_back(m_parameterDependency[k].first);
        scale.push_back(m_parameterDependency[k].second);
    }
    if (!eng->set_operators(ops4, order, scale, m_numberOfUniqueParameters, err)) {
        log_msg("Operator set-up failed: %s", err.c_str());
        return false;
    }
    // compressVector of the start (sparsematrix.h:83-105)
    std::vector<double> start(eng->dim(), 0.0);
    for (size_t i = 0; i < m_startIdx.size(); i++) {
        const long long r = eng->rank((uint32_t)m_startIdx[i]);
        if (r >= 0) start[r] = m_startCoeff[i];
    }
    if (!eng->set_start(start, err)) {
        log_msg("Start vector upload failed: %s", err.c_str());
        return false;
    }
    // Hamiltonian: triplets (src/hamiltonianmatrix.cpp:628-659) or integrals
    if (!m_iIndexes.empty()) {
        const long long N = eng->dim();
        std::vector<std::pair<long long, std::pair<int32_t, double>>> t;  // (row j, (col i, v))
        t.reserve(m_iIndexes.size());
        for (size_t q = 0; q < m_iIndexes.size(); q++) {
            if (m_iIndexes[q] < 0 || m_jIndexes[q] < 0) continue;
            const long long ci = eng->rank((uint32_t)m_iIndexes[q]);
            const long long cj = eng->rank((uint32_t)m_jIndexes[q]);
            if (ci < 0 || cj < 0) continue;
            t.push_back({cj, {(int32_t)ci, m_coeffs[q]}});
        }
        std::stable_sort(t.begin(), t.end(), [](const auto &a, const auto &b) {
            return a.first != b.first ? a.first < b.first : a.second.first < b.second.first;
        });
        std::vector<int64_t> rowptr(N + 1, 0);
        std::vector<int32_t> col;
        std::vector<double> val;
        for (size_t q = 0; q < t.size(); q++) {
            if (q > 0 && t[q].first == t[q - 1].first && t[q].second.first == t[q - 1].second.first) {
                val.back() += t[q].second.second;  // duplicates summed
                continue;
            }
            col.push_back(t[q].second.first);
            val.push_back(t[q].second.second);
            rowptr[t[q].first + 1]++;
        }
        for (long long r = 0; r < N; r++) rowptr[r + 1] += rowptr[r];
        if (!eng->set_ham_rows(rowptr, col, val, err)) {
            log_msg("Hamiltonian upload failed: %s", err.c_str());
            return false;
        }
        m_iIndexes.clear();
        m_jIndexes.clear();
        m_coeffs.clear();
    } else {
        if (2 * m_norb != m_numberOfQubits) {
            log_msg("Integrals for %d orbitals do not match %d qubits", m_norb, m_numberOfQubits);
            return false;
        }
        if (!eng->set_ham_integrals(m_norb, m_h, m_g, m_hamStorage, err)) {
            log_msg("Integral Hamiltonian set-up failed: %s", err.c_str());
            return false;
        }
    }
    m_engine = std::move(eng);
    // evolve at zero angles, as the reference does on construction
    m_angles.assign(m_excitations.size(), 0.0);
    if (!m_engine->evolve(m_angles.data(), err)) {
        log_msg("Evolution failed: %s", err.c_str());
        m_engine.reset();
        return false;
    }
    m_isConstructed = true;
    return true;
}

bool Manager::validateToRun()
{
    if (!m_isConstructed && !construct()) {
        log_msg("Not yet Constructed, something is missing");
        return false;
    }
    return true;
}

// src/AnsatzManager.cpp:467-506: per-operator angles are used as given,
// per-parameter angles are decompressed (x = D theta); psi is re-evolved
// only when the per-operator angles change.
bool Manager::setAngles(std::vector<double> angles)
{
    if (!validateToRun()) return false;
    const size_t M = m_excitations.size();
    if (angles.size() == M) {
    } else if ((int)angles.size() == m_numberOfUniqueParameters) {
        std::vector<double> dec(M);
        for (size_t k = 0; k < M; k++)
            dec[k] = m_parameterDependency[k].second * angles[m_parameterDependency[k].first];
        angles.swap(dec);
    } else {
        log_msg("Incorrect number of angles given: %zu", angles.size());
        return false;
    }
    if (angles != m_angles) {
        std::string err;
        m_angles = angles;
        if (!m_engine->evolve(m_angles.data(), err)) {
            log_msg("Evolution failed: %s", err.c_str());
            m_angles.assign(M, std::nan(""));
            return false;
        }
    }
    return true;
}

void Manager::compressGradient(const double *deriv, double *grad) const
{
    for (int p = 0; p < m_numberOfUniqueParameters; p++) grad[p] = 0;
    for (size_t k = 0; k < m_parameterDependency.size(); k++)
        grad[m_parameterDependency[k].first] += m_parameterDependency[k].second * deriv[k];
}

bool Manager::getExpectationValue(const std::vector<double> &angles, double &E)
{
    if (!setAngles(angles)) return false;
    std::string err;
    if (!m_engine->energy(&E, err)) {
        log_msg("Energy failed: %s", err.c_str());
        return false;
    }
    E += m_nuclearEnergy;
    return true;
}

bool Manager::getFinalState(const std::vector<double> &angles, std::vector<double> &full)
{
    if (!setAngles(angles)) return false;
    std::string err;
    std::vector<double> comp(m_engine->dim());
    if (!m_engine->download_state(comp.data(), err)) {
        log_msg("State download failed: %s", err.c_str());
        return false;
    }
    full.assign((size_t)1 << m_numberOfQubits, 0.0);
    const auto &st = m_engine->states();
    for (size_t i = 0; i < comp.size(); i++) full[st[i]] = comp[i];
    return true;
}

// Decompresses `angles` like setAngles; if they differ from the cached
// angles, evolves and differentiates in one fused device sequence,
// otherwise differentiates the cached psi (evolveDerivative,
// src/fusedevolve.cpp:2263-2346).
bool Manager::gradientAt(const std::vector<double> &angles, double *E, std::vector<double> &deriv)
{
    if (!validateToRun()) return false;
    const size_t M = m_excitations.size();
    std::vector<double> x;
    if (angles.size() == M) {
        x = angles;
    } else if ((int)angles.size() == m_numberOfUniqueParameters) {
        x.resize(M);
        for (size_t k = 0; k < M; k++)
            x[k] = m_parameterDependency[k].second * angles[m_parameterDependency[k].first];
    } else {
        log_msg("Incorrect number of angles given: %zu", angles.size());
        return false;
    }
    std::string err;
    deriv.resize(M);
    bool ok;
    if (x != m_angles) {
        ok = m_engine->energy_gradient(x.data(), E, deriv.data(), err);
        m_angles = ok ? x : std::vector<double>(M, std::nan(""));
    } else {
        ok = m_engine->gradient(x.data(), E, deriv.data(), err);
    }
    if (!ok) log_msg("Gradient failed: %s", err.c_str());
    return ok;
}

bool Manager::getGradientComp(const std::vector<double> &angles, std::vector<double> &grad)
{
    std::vector<double> deriv;
    if (!gradientAt(angles, nullptr, deriv)) return false;
    grad.resize(m_numberOfUniqueParameters);
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
    if (!validateToRun()) return false;
    const size_t M = m_excitations.size();
    std::vector<double> x;
    if (angles.size() == M) {
        x = angles;
    } else if ((int)angles.size() == m_numberOfUniqueParameters) {
        x.resize(M);
        for (size_t k = 0; k < M; k++)
            x[k] = m_parameterDependency[k].second * angles[m_parameterDependency[k].first];
    } else {
        log_msg("Incorrect number of angles given: %zu", angles.size());
        return false;
    }
    if (projFull.size() != ((size_t)1 << m_numberOfQubits)) {
        log_msg("Projection vector has %zu entries, expected 2^%d", projFull.size(), m_numberOfQubits);
        return false;
    }
    // the sector's amplitudes of proj (entries outside the sector meet zeros of psi)
    const auto &st = m_engine->states();
    std::vector<double> comp(st.size());
    for (size_t i = 0; i < st.size(); i++) comp[i] = projFull[st[i]];
    std::vector<double> deriv(M);
    std::string err;
    if (!m_engine->overlap_gradient(x.data(), comp.data(), &overlap, deriv.data(), err)) {
        log_msg("Overlap gradient failed: %s", err.c_str());
        m_angles.assign(M, std::nan(""));
        return false;
    }
    m_angles = x;  // psi is now U(x) psi_0
    grad.resize(m_numberOfUniqueParameters);
    compressGradient(deriv.data(), grad.data());
    return true;
}

bool Manager::getEnergyGradient(const std::vector<double> &angles, double &E, std::vector<double> &grad)
{
    std::vector<double> deriv;
    if (!gradientAt(angles, &E, deriv)) return false;
    E += m_nuclearEnergy;
    grad.resize(m_numberOfUniqueParameters);
    compressGradient(deriv.data(), grad.data());
    return true;
}

bool Manager::getHessianComp(const std::vector<double> &angles, std::vector<double> &hess)
{
    if (!setAngles(angles)) return false;
    std::string err;
    const int P = m_numberOfUniqueParameters;
    hess.assign((size_t)P * P, 0.0);
    std::vector<double> g(P);
    double E;
    if (!m_engine->hessian(m_angles.data(), hess.data(), g.data(), &E, err)) {
        log_msg("Hessian failed: %s", err.c_str());
        return false;
    }
    return true;
}

bool Manager::decompressBatch(int nvec, int nangles, const double *angles, std::vector<double> &out)
{
    const size_t M = m_excitations.size();
    out.resize((size_t)nvec * M);
    if ((size_t)nangles == M) {
        std::copy(angles, angles + (size_t)nvec * M, out.begin());
        return true;
    }
    if (nangles == m_numberOfUniqueParameters) {
        for (int v = 0; v < nvec; v++)
            for (size_t k = 0; k < M; k++)
                out[v * M + k] = m_parameterDependency[k].second *
                                 angles[(size_t)v * nangles + m_parameterDependency[k].first];
        return true;
    }
    log_msg("Incorrect number of angles given: %d", nangles);
    return false;
}

// src/AnsatzManager.cpp:709-754
bool Manager::getExpectationValues(int nvec, int nangles, const double *angles, std::vector<double> &E)
{
    if (!validateToRun()) return false;
    std::vector<double> dec;
    if (!decompressBatch(nvec, nangles, angles, dec)) return false;
    E.resize(nvec);
    std::string err;
    if (!m_engine->batch(nvec, dec.data(), E.data(), nullptr, err)) {
        log_msg("Batched energies failed: %s", err.c_str());
        return false;
    }
    for (auto &x : E) x += m_nuclearEnergy;
    return true;
}

bool Manager::getEnergyGradients(int nvec, int nangles, const double *angles, std::vector<double> &E,
                                 std::vector<double> &grads)
{
    if (!validateToRun()) return false;
    std::vector<double> dec;
    if (!decompressBatch(nvec, nangles, angles, dec)) return false;
    const size_t M = m_excitations.size();
    E.resize(nvec);
    std::vector<double> deriv((size_t)nvec * M);
    std::string err;
    if (!m_engine->batch(nvec, dec.data(), E.data(), deriv.data(), err)) {
        log_msg("Batched gradients failed: %s", err.c_str());
        return false;
    }
    grads.resize((size_t)nvec * m_numberOfUniqueParameters);
    for (int v = 0; v < nvec; v++) {
        E[v] += m_nuclearEnergy;
        compressGradient(deriv.data() + (size_t)v * M, grads.data() + (size_t)v * m_numberOfUniqueParameters);
    }
    return true;
}

bool Manager::uploadAngleBatch(int nvec, int nangles, const double *angles)
{
    if (!validateToRun()) return false;
    std::vector<double> dec;
    if (!decompressBatch(nvec, nangles, angles, dec)) return false;
    std::string err;
    if (!m_engine->batch_upload(nvec, dec.data(), err)) {
        log_msg("Batch upload failed: %s", err.c_str());
        return false;
    }
    m_batchNangles = nangles;
    return true;
}

bool Manager::runAngleBatch(bool withGradient)
{
    if (!m_isConstructed || m_engine->batch_size() < 1) {
        log_msg("No angle batch uploaded");
        return false;
    }
    std::string err;
    if (!m_engine->batch_run(withGradient, err)) {
        log_msg("Batch run failed: %s", err.c_str());
        return false;
    }
    return true;
}

bool Manager::downloadAngleBatch(double *energies, double *gradients)
{
    if (!m_isConstructed || m_engine->batch_size() < 1) {
        log_msg("No angle batch uploaded");
        return false;
    }
    const int nvec = m_engine->batch_size();
    const size_t M = m_excitations.size();
    std::vector<double> deriv(gradients ? (size_t)nvec * M : 0);
    std::string err;
    if (!m_engine->batch_download(energies, gradients ? deriv.data() : nullptr, err)) {
        log_msg("Batch download failed: %s", err.c_str());
        return false;
    }
    for (int v = 0; v < nvec; v++) {
        energies[v] += m_nuclearEnergy;
        if (gradients)
            compressGradient(deriv.data() + (size_t)v * M, gradients + (size_t)v * m_numberOfUniqueParameters);
    }
    return true;
}

bool Manager::setKernelTiming(bool on)
{
    if (!validateToRun()) return false;
    m_engine->set_timing(on);
    return true;
}

bool Manager::getKernelTiming(int cls, double &ms, long long &launches)
{
    if (!validateToRun()) return false;
    std::string err;
    if (!m_engine->timing(cls, &ms, &launches, err)) {
        log_msg("Timing query failed: %s", err.c_str());
        return false;
    }
    return true;
}

bool Manager::getDimensions(long long &dim, int &nops, int &nparams)
{
    if (!validateToRun()) return false;
    dim = m_engine->dim();
    nops = (int)m_excitations.size();
    nparams = m_numberOfUniqueParameters;
    return true;
}

} // namespace ae

namespace ae {

bool Manager::optimise(std::vector<double> &angles, double &E, std::vector<double> &trace, double stopEnergy,
                       bool useStop)
{
    if (!validateToRun()) return false;
    const size_t M = m_excitations.size();
    const int P = m_numberOfUniqueParameters;
    const bool perParam = (int)angles.size() == P && angles.size() != M;
    std::vector<double> x;
    if (angles.size() == M) x = angles;
    else if (perParam) {
        x.resize(M);
        for (size_t k = 0; k < M; k++)
            x[k] = m_parameterDependency[k].second * angles[m_parameterDependency[k].first];
    } else {
        log_msg("Incorrect number of angles given: %zu", angles.size());
        return false;
    }
    std::vector<int> order(M);
    std::vector<double> scale(M);
    for (size_t k = 0; k < M; k++) {
        order[k] = m_parameterDependency[k].first;
        scale[k] = m_parameterDependency[k].second;
    }
    Engine *eng = m_engine.get();
    std::string err;
    EvalFns f;
    f.hessian = [&](const std::vector<double> &a, std::vector<double> &H, std::vector<double> &g, double &En) {
        if (!eng->hessian(a.data(), H.data(), g.data(), &En, err)) return false;
        En += m_nuclearEnergy;
        return true;
    };
    f.egrad = [&](const std::vector<double> &a, double &En, std::vector<double> &g) {
        std::vector<double> deriv(M);
        if (!eng->evolve(a.data(), err) || !eng->gradient(a.data(), &En, deriv.data(), err)) return false;
        En += m_nuclearEnergy;
        compressGradient(deriv.data(), g.data());
        return true;
    };
    NewtonDriver nd(order, scale, P, f);
    if (useStop) nd.set_stop_energy(stopEnergy);
    const bool ok = nd.optimise(x, E);
    // psi now belongs to the last evaluated angles; force a re-evolve next time
    m_angles.assign(M, std::nan(""));
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
            const int p = m_parameterDependency[k].first;
            if (!seen[p]) {
                angles[p] = x[k] / m_parameterDependency[k].second;
                seen[p] = true;
            }
        }
    } else {
        angles = x;
    }
    return true;
}

} // namespace ae
