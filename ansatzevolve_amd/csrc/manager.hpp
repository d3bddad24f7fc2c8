// Host-side counterpart of the reference's stateAnsatzManager
// (src/AnsatzManager.{h,cpp}), with the same public operations: it records
// the set-up, builds the device engine on the first evaluation, picks the
// sector from the conserved quantities of the start state and the
// operators, converts angle formats and remembers the angles of the current
// state.  The evaluations run on the device engine.
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
    std::mutex lock;  // taken by every C-ABI call (the reference's per-manager mutex, src/AnsatzManager.h:121)

    bool storeOperators(const std::vector<std::array<int8_t, 4>> &excs);
    bool storeInitial(int numberOfQubits, const std::vector<int> &indexes, const std::vector<double> &coeffs);
    bool storeInitialComplex(int numberOfQubits, const std::vector<int> &indexes, const std::vector<double> &re,
                             const std::vector<double> &im);
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
    // the real part for complex amplitudes (src/AnsatzSynthInterface.cpp:402-407)
    bool getFinalState(const std::vector<double> &angles, std::vector<double> &full);
    // both parts over the full space (imaginary part zero for real amplitudes)
    bool getFinalStateComplex(const std::vector<double> &angles, std::vector<double> &re, std::vector<double> &im);
    // Complex amplitudes: the start state was given with complex
    // coefficients or some operator is a number operator.
    bool complexAmplitudes() const;
    bool getGradientComp(const std::vector<double> &angles, std::vector<double> &grad);
    // per operator (decompressed; src/AnsatzManager.h:221)
    bool getGradient(const std::vector<double> &angles, std::vector<double> &grad);
    // overlap <proj|psi(angles)> and its per-parameter gradient, proj over
    // the full space of 2^qubits basis states (src/fusedevolve.cpp:2814-2889)
    bool getOverlapGradient(const std::vector<double> &angles, const std::vector<double> &projFull,
                            double &overlap, std::vector<double> &grad);
    bool getHessianComp(const std::vector<double> &angles, std::vector<double> &hessRowMajor);
    // Hessian (row-major), gradient and value of the overlap <proj|psi>, proj
    // over the full space of 2^qubits basis states (src/fusedevolve.cpp:2892-3183)
    bool getHessianProj(const std::vector<double> &angles, const std::vector<double> &projFull,
                        std::vector<double> &hessRowMajor, std::vector<double> &grad, double &overlap);
    bool getEnergyGradient(const std::vector<double> &angles, double &E, std::vector<double> &grad);
    // getEnergyGradient without any allocation (Engine::energy_gradient_noalloc).
    // 0 done, 1 not built yet (building allocates), 2 the set-up needs more
    // scratch than the fixed array holds, 3 other failure.
    int getEnergyGradientNoAlloc(const double *angles, int nangles, double &E, double *grad);
    bool built() const { return eng_ != nullptr; }
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
    // The Hamiltonian in the span of the states of nsets angle vectors
    // (doSubspaceDiagonalisation): Loewdin-symmetrised eigenvalues, the
    // Gram matrices (row-major) and the number of dropped directions.
    bool subspaceDiagonalise(int nsets, int nangles, const double *angles, std::vector<double> &evals,
                             std::vector<double> &H, std::vector<double> &S, int &nullRank);
    // writeProperties: one row of kPropColumns + 2 noverlaps values per
    // angle vector (see the C interface), the Hessians and metric tensors
    // (row-major P x P each, unless noHess) and the lowest noverlaps
    // eigenvalues of H (when noverlaps > 0).
    static constexpr int kPropColumns = 12;
    bool properties(int nsets, int nangles, const double *angles, int noverlaps, bool noHess,
                    std::vector<double> &rows, std::vector<double> &hessians, std::vector<double> &metrics,
                    std::vector<double> &lowest);
    // The iterative projected optimiser (projected.cpp); angles and trace as optimise.
    bool iterativeOptimise(std::vector<double> &angles, double &E, std::vector<double> &trace, bool ediis);
    bool getKernelTiming(int cls, double &ms, long long &launches);
    // Execution choices of the device engine (EngineOption) and the
    // autotuner that picks them for a batch shape (autotune.cpp).
    bool setEngineOption(int key, int value);
    bool getEngineOption(int key, int &value);
    bool autotune(int nvec, int nangles, const double *angles, const char *cacheFile, int reps, std::string &report);

    int numberOfParameters() const { return nparams_; }
    int batchAngleCount() const { return batchNangles_; }
    // The device engine once built (nullptr before); used by the
    // multi-instance batch, which holds every involved manager's lock.
    Engine *engine() { return eng_.get(); }

private:
    // Builds the device engine on the first evaluation; false (with the
    // reason logged) while something the evaluation needs is missing.
    bool ensureBuilt();
    bool build();
    // True, with a log line naming `what`, once the engine is built: the
    // set-up is fixed from the first evaluation on.
    bool frozen(const char *what) const;
    // The sector the engine works in: S_z when the start state has a fixed
    // (up, down) occupation and every operator conserves it; particle number
    // when only the total is fixed; else the full space.
    int chooseSector() const;
    // Per-operator angles from angles given per operator (M) or per
    // parameter (P, x_k = scale_k theta_order_k); false for other lengths.
    bool toOperatorAngles(const double *in, size_t n, double *out) const;
    bool toOperatorAngles(const std::vector<double> &in, std::vector<double> &out) const;
    bool setAngles(std::vector<double> angles);
    bool gradientAt(const std::vector<double> &angles, double *E, std::vector<double> &deriv);
    // True for real amplitudes; else logs that `what` needs them.
    bool realOnly(const char *what) const;
    bool storeStart(int numberOfQubits, const std::vector<int> &indexes, const std::vector<double> &coeffs,
                    const std::vector<double> *imag);
    bool decompressBatch(int nvec, int nangles, const double *angles, std::vector<double> &out);
    void compressGradient(const double *deriv, double *grad) const;

    // --- the stored set-up
    struct Start {
        int qubits = -1;
        std::vector<int> bits;      // zero-based basis states
        std::vector<double> amps;
        std::vector<double> ampsIm;  // imaginary parts (setInitialStateComplex)
        bool complex = false;        // given with complex coefficients
    } start_;
    // occupation shared by every nonzero component of the start state
    struct Occupation {
        bool fixedTotal = false;  // every component has `total` electrons
        bool fixedSpins = false;  // ... and the same (up, down) split
        int total = -1, up = -1, down = -1;
    } occ_;
    std::vector<std::array<int8_t, 4>> ops_;  // zero-based, -1 absent
    std::vector<int> paramIdx_;               // parameter of each operator
    std::vector<double> paramScale_;          // its scale
    int nparams_ = -1;
    struct Triplets {
        std::vector<int> i, j;  // zero-based full-space indices
        std::vector<double> v;
    } trip_;
    int norb_ = 0;
    std::vector<double> h1_, g2_;  // one- and two-electron integrals
    int storage_ = 0;               // HamStorage
    std::string runPrefix_;
    double enuc_ = 0;

    // --- the built state
    std::unique_ptr<Engine> eng_;
    std::vector<double> cur_;  // per-operator angles of the engine's current psi
    int batchNangles_ = 0;     // angle count of the uploaded batch
    std::vector<double> nostage_, noderiv_;  // host staging of the no-allocation evaluation
};

} // namespace ae
