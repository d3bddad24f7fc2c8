// Host-side mirror of the reference's stateAnsatzManager
// (src/AnsatzManager.{h,cpp}): stores the set-up, constructs lazily on the
// first evaluation, detects the particle-number / S_z symmetry that picks
// the sector, converts angle formats and caches the last angles.  The
// evaluations themselves run on the device engine.
#pragma once
#include <array>
#include <memory>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "engine.hpp"
#include "newton.hpp"

namespace ae {

class Manager {
public:
    std::mutex lock;  // taken by every C-ABI call (src/AnsatzManager.h:121)

    bool storeOperators(const std::vector<std::array<int8_t, 4>> &excs);
    bool storeInitial(int numberOfQubits, const std::vector<int> &indexes, const std::vector<double> &coeffs);
    bool storeHamiltonian(std::vector<int> &&i0, std::vector<int> &&j0, std::vector<double> &&v);
    bool storeIntegrals(int norb, const double *h, const double *g);
    bool storeNuclearEnergy(double e);
    bool storeParameterDependencies(const std::vector<std::pair<int, double>> &dep);
    bool storeHamiltonianStorage(int mode);
    // Run path prefix (src/AnsatzManager.cpp:447-466): with no Hamiltonian
    // stored, construction loads it from <prefix>_Ham_Coeff/_Ham_Index.dat
    // or <prefix>_oneEInts/_twoEInts.bin (src/AnsatzManager.cpp:185-199).
    bool storeRunPath(const std::string &runPath);

    bool getExpectationValue(const std::vector<double> &angles, double &E);
    bool getFinalState(const std::vector<double> &angles, std::vector<double> &full);
    bool getGradientComp(const std::vector<double> &angles, std::vector<double> &grad);
    // per operator (decompressed; src/AnsatzManager.h:221)
    bool getGradient(const std::vector<double> &angles, std::vector<double> &grad);
    // overlap <proj|psi(angles)> and its per-parameter gradient, proj over
    // the full space of 2^qubits basis states (src/fusedevolve.cpp:2814-2889)
    bool getOverlapGradient(const std::vector<double> &angles, const std::vector<double> &projFull,
                            double &overlap, std::vector<double> &grad);
    bool getHessianComp(const std::vector<double> &angles, std::vector<double> &hessRowMajor);
    bool getEnergyGradient(const std::vector<double> &angles, double &E, std::vector<double> &grad);
    bool getExpectationValues(int nvec, int nangles, const double *angles, std::vector<double> &E);
    bool getEnergyGradients(int nvec, int nangles, const double *angles, std::vector<double> &E,
                            std::vector<double> &grads);
    bool getDimensions(long long &dim, int &nops, int &nparams);

    bool uploadAngleBatch(int nvec, int nangles, const double *angles);
    bool runAngleBatch(bool withGradient);
    bool downloadAngleBatch(double *energies, double *gradients);
    bool setKernelTiming(bool on);
    // Plain Newton optimiser from `angles` (per operator or per parameter,
    // returned in the same form); trace rows: seconds, energy, |g|,
    // Hessian evaluations, energy evaluations.
    bool optimise(std::vector<double> &angles, double &E, std::vector<double> &trace, double stopEnergy,
                  bool useStop);
    bool getKernelTiming(int cls, double &ms, long long &launches);
    // Execution choices of the device engine (EngineOption) and the
    // autotuner that picks them for a batch shape (autotune.cpp).
    bool setEngineOption(int key, int value);
    bool getEngineOption(int key, int &value);
    bool autotune(int nvec, int nangles, const double *angles, const char *cacheFile, int reps, std::string &report);

    int numberOfParameters() const { return m_numberOfUniqueParameters; }
    int batchAngleCount() const { return m_batchNangles; }
    // The device engine after construction (nullptr before); used by the
    // multi-instance batch, which holds every involved manager's lock.
    Engine *engine() { return m_isConstructed ? m_engine.get() : nullptr; }

private:
    bool validateToRun();
    bool construct();
    void setOperatorSymmetry();
    bool setAngles(std::vector<double> angles);
    bool gradientAt(const std::vector<double> &angles, double *E, std::vector<double> &deriv);
    bool decompressBatch(int nvec, int nangles, const double *angles, std::vector<double> &out);
    void compressGradient(const double *deriv, double *grad) const;

    bool m_isConstructed = false;
    double m_nuclearEnergy = 0;
    int m_numberOfUniqueParameters = -1;
    std::vector<std::array<int8_t, 4>> m_excitations;
    std::vector<std::pair<int, double>> m_parameterDependency;
    std::vector<int> m_startIdx;
    std::vector<double> m_startCoeff;
    int m_numberOfQubits = -1;
    int m_numberOfParticles = -1, m_spinUp = -1, m_spinDown = -1;
    bool m_InitialSZSym = false, m_OperatorSZSym = false, m_particleSym = false;
    std::vector<int> m_iIndexes, m_jIndexes;
    std::vector<double> m_coeffs;
    int m_norb = 0;
    std::vector<double> m_h, m_g;
    int m_hamStorage = 0;
    std::string m_runPath;

    std::unique_ptr<Engine> m_engine;
    std::vector<double> m_angles;  // per-operator angles of the current psi
    int m_batchNangles = 0;        // angle count of the uploaded batch
};

} // namespace ae
