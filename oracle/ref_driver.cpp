// Test infrastructure only: a thin C driver over the reference's own C++
// classes, compiled together with the reference sources where they lie (see
// oracle/Makefile).  Used to produce golden vectors and to time the reference
// CPU path.  Nothing in the product links or loads this.
//
// Two routes into the reference:
//  * refd_* "manager" calls go through stateAnsatzManager exactly as the
//    reference C-ABI does (src/AnsatzSynthInterface.cpp:313-629), but accept an
//    integral-file prefix so the integral Hamiltonian path
//    (src/hamiltonianmatrix.cpp:661-706) is reachable without 1e9 triplets.
//  * refd_fe_* calls build FusedEvolve directly, the way the reference CLI
//    benchmark mode does (src/main.cpp:404-435), so that one "E+grad" is
//    FusedEvolve::evolve + FusedEvolve::evolveDerivative(&E).
#include "AnsatzManager.h"
#include "fusedevolve.h"
#include "hamiltonianmatrix.h"
#include "operatorpool.h"
#include "tupsquantities.h"
#include "threadpool.h"

#include <chrono>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

namespace {

std::vector<stateRotate::exc> make_excs(int nops, const int* ops)
{
    std::vector<stateRotate::exc> excs(nops);
    for (int i = 0; i < nops; i++)
        for (int j = 0; j < 4; j++)
            excs[i][j] = (int8_t)ops[4 * i + j];
    return excs;
}

std::vector<std::pair<int, realNumType>> make_order(int nops, const int* order0, const double* scale)
{
    std::vector<std::pair<int, realNumType>> dep;
    for (int i = 0; i < nops; i++)
        dep.push_back({order0[i], scale ? scale[i] : 1.0});
    return dep;
}

struct FEHandle {
    std::shared_ptr<compressor> comp;
    std::shared_ptr<HamiltonianMatrix<realNumType, numType>> ham;
    std::shared_ptr<TUPSQuantities> tq;
    std::shared_ptr<FusedEvolve> fe;
    vector<numType> dest;
    int nparams = 0;
    int nops = 0;
};

} // namespace

extern "C" {

// Manager route.  ops: nops x 4 zero-based spin-orbital indices (-1 = absent);
// order0: zero-based parameter index per operator; init_idx0: zero-based
// basis indices.  Hamiltonian: either triplets (0-based full-space indices)
// or, when ntrip == 0, integral files at `prefix` (_oneEInts.bin/_twoEInts.bin).
void* refd_create(int nq, int ninit, const int* init_idx0, const double* init_coeff,
                  int nops, const int* ops, const int* order0, const double* scale,
                  const char* prefix, int ntrip, const int* ti0, const int* tj0, const double* tv)
{
    auto* m = new stateAnsatzManager;
    std::vector<int> idx(init_idx0, init_idx0 + ninit);
    std::vector<numType> co(init_coeff, init_coeff + ninit);
    bool ok = m->storeInitial(nq, idx, co);
    if (ntrip > 0) {
        ok = ok && m->storeHamiltonian(std::vector<int>(ti0, ti0 + ntrip), std::vector<int>(tj0, tj0 + ntrip),
                                       std::vector<realNumType>(tv, tv + ntrip));
    } else {
        ok = ok && m->storeRunPath(std::string(prefix));
    }
    ok = ok && m->storeOperators(make_excs(nops, ops));
    ok = ok && m->storeParameterDependencies(make_order(nops, order0, scale));
    if (!ok) {
        delete m;
        return nullptr;
    }
    return m;
}

void refd_destroy(void* h) { delete static_cast<stateAnsatzManager*>(h); }

int refd_energy(void* h, int na, const double* angles, double* E)
{
    auto* m = static_cast<stateAnsatzManager*>(h);
    realNumType e = 0;
    if (!m->getExpectationValue(std::vector<double>(angles, angles + na), e)) return 1;
    *E = e;
    return 0;
}

int refd_gradient(void* h, int na, const double* angles, double* g, int ng)
{
    auto* m = static_cast<stateAnsatzManager*>(h);
    vector<realNumType> gv;
    if (!m->getGradientComp(std::vector<double>(angles, angles + na), gv)) return 1;
    if ((int)gv.size() != ng) return 2;
    std::memcpy(g, &gv[0], sizeof(double) * ng);
    return 0;
}

// Column-major P x P output, as the reference C-ABI writes it.
int refd_hessian(void* h, int na, const double* angles, double* H, int np)
{
    auto* m = static_cast<stateAnsatzManager*>(h);
    Matrix<realNumType>::EigenMatrix hm;
    if (!m->getHessianComp(std::vector<double>(angles, angles + na), hm)) return 1;
    if (hm.rows() != np || hm.cols() != np) return 2;
    for (int i = 0; i < np; i++)
        for (int j = 0; j < np; j++) H[np * j + i] = hm(i, j);
    return 0;
}

int refd_state(void* h, int na, const double* angles, double* out, long nfull)
{
    auto* m = static_cast<stateAnsatzManager*>(h);
    vector<numType> st;
    if (!m->getFinalState(std::vector<double>(angles, angles + na), st)) return 1;
    if ((long)st.size() != nfull) return 2;
    std::memcpy(out, &st[0], sizeof(double) * nfull);
    return 0;
}

// FusedEvolve route (integral files only, Sz-and-number compressed sector).
void* refd_fe_create(int nq, int spin_up, int spin_down, int ninit, const int* init_idx0, const double* init_coeff,
                     int nops, const int* ops, const int* order0, const double* scale, int nparams,
                     const char* prefix)
{
    threadpool::getInstance(NUM_CORES);
    auto* H = new FEHandle;
    H->comp = std::make_shared<SZAndnumberOperatorCompressor>(1u << nq, spin_up, spin_down);
    vector<numType> full;
    full.resize(1ul << nq, false, nullptr);
    for (int i = 0; i < ninit; i++) full[init_idx0[i]] = init_coeff[i];
    vector<numType> start;
    compressor::compressVector<numType>(full, start, H->comp);
    H->ham = std::make_shared<HamiltonianMatrix<realNumType, numType>>(std::string(prefix), nq, H->comp);
    if (!H->ham->ok()) {
        delete H;
        return nullptr;
    }
    H->tq = std::make_shared<TUPSQuantities>(H->ham, make_order(nops, order0, scale), nparams, 0.0, std::string(prefix));
    H->fe = std::make_shared<FusedEvolve>(start, H->ham, H->tq->m_compressMatrix, H->tq->m_deCompressMatrix);
    H->fe->updateExc(make_excs(nops, ops));
    H->nparams = nparams;
    H->nops = nops;
    return H;
}

void refd_fe_destroy(void* h) { delete static_cast<FEHandle*>(h); }

// One E+grad at per-operator angles (length nops): evolve + evolveDerivative(&E).
// Returns wall seconds of the last of `reps` repetitions' mean.
double refd_fe_egrad(void* h, const double* angles_ops, int reps, double* E, double* grad_comp)
{
    auto* H = static_cast<FEHandle*>(h);
    std::vector<realNumType> ang(angles_ops, angles_ops + H->nops);
    vector<realNumType> deriv;
    realNumType e = 0;
    auto t0 = std::chrono::high_resolution_clock::now();
    for (int r = 0; r < reps; r++) {
        H->fe->evolve(H->dest, ang);
        H->fe->evolveDerivative(H->dest, deriv, ang, &e);
    }
    auto t1 = std::chrono::high_resolution_clock::now();
    *E = e;
    for (int i = 0; i < H->nparams; i++) grad_comp[i] = deriv[i];
    return std::chrono::duration<double>(t1 - t0).count() / reps;
}

// FusedEvolve::evolveHessianProj at per-operator angles (length nops), proj
// over the full space of 2^nq states; column-major P x P Hessian.
int refd_fe_hessian_proj(void* h, const double* angles_ops, const double* proj_full, int nq, double* hess,
                         double* grad_comp, double* ov)
{
    auto* H = static_cast<FEHandle*>(h);
    std::vector<realNumType> ang(angles_ops, angles_ops + H->nops);
    vector<numType> full;
    full.resize(1ul << nq, false, nullptr);
    for (long i = 0; i < (1l << nq); i++) full[i] = proj_full[i];
    vector<numType> proj;
    compressor::compressVector<numType>(full, proj, H->comp);
    Eigen::MatrixXd hm;
    vector<realNumType> g;
    realNumType e = 0;
    H->fe->evolveHessianProj(hm, g, ang, proj, nullptr, &e);
    if (hm.rows() != H->nparams || hm.cols() != H->nparams) return 2;
    for (int i = 0; i < H->nparams; i++)
        for (int j = 0; j < H->nparams; j++) hess[H->nparams * j + i] = hm(i, j);
    for (int i = 0; i < H->nparams; i++) grad_comp[i] = g[i];
    *ov = e;
    return 0;
}

// TUPSQuantities::iterativeTups on the FusedEvolve route from per-operator
// angles (updated in place); returns its final energy.
double refd_fe_iterative(void* h, double* angles_ops)
{
    auto* H = static_cast<FEHandle*>(h);
    std::vector<baseAnsatz::rotationElement> rp(H->nops);
    for (int i = 0; i < H->nops; i++) rp[i] = {0, angles_ops[i]};
    const double E = H->tq->iterativeTups(*H->fe, rp, true);
    for (int i = 0; i < H->nops; i++) angles_ops[i] = rp[i].second;
    return E;
}

long refd_fe_dim(void* h) { return (long)static_cast<FEHandle*>(h)->comp->getCompressedSize(); }

int refd_num_threads(void) { return (int)NUM_CORES; }

} // extern "C"
