// Test infrastructure only: a thin C driver over the reference's own C++
// classes compiled in the reference's complex mode (its COMPLEX_MODE build
// option, the useComplex define; see oracle/Makefile target ref-complex).
// Used to write golden vectors for number-operator (diagonal) rotations and
// complex start states.  Nothing in the product links or loads this.
#include "AnsatzManager.h"

#include <complex>
#include <cstring>
#include <vector>

extern "C" {

// Manager route, as the reference's C interface drives it
// (src/AnsatzSynthInterface.cpp): complex start state (re, im arrays),
// zero-based full-space triplets, operators nops x 4 (-1 absent), order.
void* refc_create(int nq, int ninit, const int* init_idx0, const double* init_re, const double* init_im, int nops,
                  const int* ops, const int* order0, const double* scale, int ntrip, const int* ti0, const int* tj0,
                  const double* tv)
{
    auto* m = new stateAnsatzManager;
    std::vector<int> idx(init_idx0, init_idx0 + ninit);
    std::vector<std::complex<double>> co(ninit);
    for (int i = 0; i < ninit; i++) co[i] = std::complex<double>(init_re[i], init_im[i]);
    bool ok = m->storeInitial(nq, idx, co);
    ok = ok && m->storeHamiltonian(std::vector<int>(ti0, ti0 + ntrip), std::vector<int>(tj0, tj0 + ntrip),
                                   std::vector<realNumType>(tv, tv + ntrip));
    std::vector<stateRotate::exc> excs(nops);
    for (int i = 0; i < nops; i++)
        for (int j = 0; j < 4; j++) excs[i][j] = (int8_t)ops[4 * i + j];
    ok = ok && m->storeOperators(excs);
    std::vector<std::pair<int, realNumType>> dep;
    for (int i = 0; i < nops; i++) dep.push_back({order0[i], scale ? scale[i] : 1.0});
    ok = ok && m->storeParameterDependencies(dep);
    if (!ok) {
        delete m;
        return nullptr;
    }
    return m;
}

void refc_destroy(void* h) { delete static_cast<stateAnsatzManager*>(h); }

int refc_energy(void* h, int na, const double* angles, double* E)
{
    auto* m = static_cast<stateAnsatzManager*>(h);
    realNumType e = 0;
    if (!m->getExpectationValue(std::vector<double>(angles, angles + na), e)) return 1;
    *E = e;
    return 0;
}

int refc_gradient(void* h, int na, const double* angles, double* g, int ng)
{
    auto* m = static_cast<stateAnsatzManager*>(h);
    vector<realNumType> gv;
    if (!m->getGradientComp(std::vector<double>(angles, angles + na), gv)) return 1;
    if ((int)gv.size() != ng) return 2;
    std::memcpy(g, &gv[0], sizeof(double) * ng);
    return 0;
}

// final state over the full space: re and im arrays of nfull entries
int refc_state(void* h, int na, const double* angles, double* re, double* im, long nfull)
{
    auto* m = static_cast<stateAnsatzManager*>(h);
    vector<numType> st;
    if (!m->getFinalState(std::vector<double>(angles, angles + na), st)) return 1;
    if ((long)st.size() != nfull) return 2;
    for (long i = 0; i < nfull; i++) {
        re[i] = std::real(st[i]);
        im[i] = std::imag(st[i]);
    }
    return 0;
}

} // extern "C"
