// Device engine: the statevector, the sector index maps, the operator
// sweeps and the Hamiltonian, all resident in HBM, driven on one HIP stream.
#pragma once
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "common.hpp"

namespace ae {

struct EngineImpl;

// Which form the Hamiltonian takes on the device.
enum HamStorage : int { kHamAuto = 0, kHamExplicit = 1, kHamMatrixFree = 2 };

// How operator sweeps are fused (S_z sector).
//   kSweepOrbit:    runs of operators touching <= 2 alpha and <= 2 beta
//                   orbitals as register orbits, the rest as LDS tiles;
//   kSweepOrbitLds: the same, orbit segments staged through LDS rows;
//   kSweepTile:     LDS tiles only;
//   kSweepOrbit2:   two TUPS units on disjoint orbital pairs per pass,
//                   thread per orbit (k_orb2s), when every operator falls
//                   into such segments (else as kSweepOrbit).
// Batches in the lane layout use their own segmentation (kOptLaneGroups).
enum SweepKind : int { kSweepOrbit = 0, kSweepOrbitLds = 1, kSweepTile = 2, kSweepOrbit2 = 3 };

// Execution choices (Engine::set_option); none changes a result beyond
// floating-point summation order.
enum EngineOption : int {
    kOptSweep = 0,    // SweepKind
    kOptLaneMin = 1,  // batches of >= value vectors use the lane layout; 0 never
    kOptGraph = 2,    // 1: replay hipGraphs, 0: eager launches
    kOptFuseDepth = 3,  // most operators fused into one pass (0: no limit beyond the kernel's)
    kOptLdsBudget = 4,  // bytes of LDS an LDS-tile segment may use (0: the device's limit)
    kOptHamKind = 5,    // read only: 1 rows, 2 uniform-width rows, 3 matrix-free
    kOptSegments = 6,   // read only: fused segments per sweep (vector-major sweeps)
    kOptLaneGroups = 7, // lane-layout adjoint sweep: disjoint 2-bit orbital groups per spin and pass (1 to 3)
    kOptLaneSegments = 8,  // read only: fused segments per lane-layout adjoint sweep
    kOptLaneGroupsFwd = 9, // the same for the lane-layout forward sweep
    kOptCheckpoint = 10,   // 1: the lane-layout gradient keeps every segment's forward state (out-of-place
                           // forward, adjoint without state write-back) when the two sweeps share their
                           // segments and the states fit in half the free memory; 0: in-place sweeps
    kOptCheckpointActive = 11,  // read only: 1 when the uploaded batch runs with stored states
    kOptSigmaHalves = 12,       // 1: lane-layout sigma with double-buffered source halves when available,
                                // 0 (default): the one-row kernel
    kOptSigmaHalvesAvailable = 13,  // read only: 1 when the two-half sigma applies (n = 2 kd <= 12)
    kOptSigmaGrouped = 14,      // 1 (default): lane-layout sigma with targets grouped by their upper six
                                // beta orbitals when available (takes precedence over kOptSigmaHalves)
    kOptSigmaGroupedAvailable = 15,  // read only: 1 when the grouped sigma applies (12 orbitals, 6 beta)
    kOptGemmSolutionBeta = 16,   // rocBLAS solution index of the lane sigma's beta-part DGEMM (0: default)
    kOptGemmSolutionAlpha = 17,  // the same for the alpha-part DGEMM (an index the library lacks: default)
    kOptReachable = 18,          // 1 (default): lane-layout sweeps skip orbits that the start state and the
                                 // operators before them cannot reach (exact zeros); 0: every orbit
    // read only, after a lane-layout upload: amplitudes per vector and step (summed over the segments)
    kOptSweepAmpForward = 19,        // read and written by the forward sweep
    kOptSweepAmpAdjointLambda = 20,  // of lambda read and written by the adjoint sweep
    kOptSweepAmpAdjointState = 21,   // of the forward state read by the adjoint sweep
    kOptSweepTwoGroupAvailable = 22,  // read only: 1 when kSweepOrbit2 has two-group segments to run
    kOptSigma1Rows = 23,  // vector-major split sigma: source rows per phase (2, 4, 6 or 8 (default); 0: the two-row kernel);
                          // a count whose LDS exceeds the device's falls back to the largest that fits
    kOptVectorSegments = 24,  // read only: segments per vector-major sweep (launches, the two-group list with kind 3)
    kOptWindow = 25,          // lane-layout sweeps as window passes over w consecutive spatial orbitals (2 to 8;
                              // 0: orbit segments; -1, the default: w = 4 when the tiles fit in registers, else
                              // orbit segments); used when every operator falls into a window phase
    kOptWindowWaves = 26,     // waves per block of the window-pass kernels: 4 (default) or 8
    kOptWindowPasses = 27,    // read only: window passes per sweep (0: window passes not in use)
    kOptWindowRegisters = 28, // read only: 1 when the window passes hold their tiles in registers (four-orbital
                              // windows, every phase on one pair in both spins), 2 when every phase is also a
                              // five-operator TUPS unit (straight-line phases), 0 otherwise
    kOptWindowVector = 29,    // window passes for the vector-major sweeps (single vectors, small batches): -1, the
                              // default, from 2^18 determinants on; 0 never; 1 at any size. They apply only when
                              // the window passes hold register tiles (key 28 is 1 or 2)
    kOptWindowEnergy = 30,    // 1 (default): batches with the gradient and register window tiles take the energies
                              // in the first adjoint pass instead of a separate dot; 0: the separate dot
    kOptWindowHessian = 31,   // 1 (default): the Hessian's tangent sweep runs as window passes when the window
                              // passes hold TUPS units in registers (key 28 = 2); 0: two-group segments
    kOptWindowStaged = 32,    // 1 (default): the vector-major window passes stage each alpha class's rows in LDS
                              // (k_winvs) where they fit; 0: one tile per lane straight from memory (k_winv).
                              // Read back: 1 in use, 2 requested but the rows do not fit, 0 off
    kOptSigmaExchange = 33,   // -1 (default): a lane-layout batch whose states are symmetric under the exchange
                              // of the alpha and beta strings (equal strings per spin, the same same-spin matrix,
                              // an exchange-symmetric start vector, every operator its own mirror or next to its
                              // mirror with the same angle in every vector) takes the beta-beta part of the
                              // grouped sigma from the alpha-alpha DGEMM's result read transposed, instead of a
                              // second DGEMM; 0: never.  Read back: 1 in use for the uploaded batch, 2 allowed
                              // but not applicable, 0 off
};

class Engine {
public:
    Engine();
    ~Engine();
    Engine(const Engine &) = delete;
    Engine &operator=(const Engine &) = delete;

    // Sector: nq qubits; mode kSz uses (nup, ndown), kNumber uses nel.
    bool init(int nq, int mode, int nel, int nup, int ndown, std::string &err);
    long long dim() const;
    int mode() const;
    // Host copy of the sector's bit strings, ascending.
    const std::vector<uint32_t> &states() const;
    // Sector index of a full-space bit string, -1 if outside.
    long long rank(uint32_t bits) const;

    // Operators in circuit order (M x 4, zero-based, -1 absent) and the
    // parameter map (order[k] = parameter of operator k, scale[k]).
    bool set_operators(const std::vector<int8_t> &ops4, const std::vector<int> &order,
                       const std::vector<double> &scale, int nparams, std::string &err);
    // psi_0 over the sector.
    bool set_start(const std::vector<double> &start, std::string &err);

    // Complex amplitudes (the reference's complex build, useComplex in
    // src/globals.h:16-28): psi is held as a real and an imaginary plane,
    // the excitations rotate both alike, and number operators (a^+_p a_p,
    // a^+_p a^+_q a_p a_q) multiply the amplitudes of the states where they
    // act by exp(+i angle) (src/fusedevolve.cpp:520-717).  Evolution,
    // energy, gradient and the state are available; the Hessians, overlap
    // gradients, batches, subspace products and the no-allocation
    // evaluation refuse.  Switching re-segments the operators and drops the
    // start state.  set_start gives a real start (imaginary plane zero).
    bool set_complex(bool on, std::string &err);
    bool is_complex() const;
    bool set_start_complex(const std::vector<double> &re, const std::vector<double> &im, std::string &err);
    // The current psi over the sector as its two planes (im all zero for
    // real amplitudes).
    bool download_state_complex(double *re, double *im, std::string &err);

    // Hamiltonian as rows of y = M^T x over the sector (output row j holds
    // (input index, value) pairs), built on the host from triplets.
    bool set_ham_rows(const std::vector<int64_t> &rowptr, const std::vector<int32_t> &col,
                      const std::vector<double> &val, std::string &err);
    // Hamiltonian from integrals (S_z sector only), built on the device.
    bool set_ham_integrals(int norb, const std::vector<double> &h, const std::vector<double> &g, int storage,
                           std::string &err);

    // psi = U(angles) psi_0; angles per operator.
    bool evolve(const double *angles, std::string &err);
    // E = <psi|H|psi> for the current psi.
    bool energy(double *E, std::string &err);
    // Adjoint sweep from the current psi (which must be U(angles) psi_0):
    // deriv[k] = dE/d(angle_k) per operator; E optional.
    bool gradient(const double *angles, double *E, double *deriv, std::string &err);
    // psi = U(angles) psi_0, then E and the adjoint sweep, as one device
    // sequence (replayed from a hipGraph after the first call).
    bool energy_gradient(const double *angles, double *E, double *deriv, std::string &err);
    // The same without allocating: no graph, no library call, no device
    // allocation; the angle tables, partial sums and results live in a
    // fixed device array shared by every call on the GPU (calls that
    // overlap on one GPU race).  Fails when the set-up needs more than
    // noalloc_scratch_doubles() doubles of it (noalloc_scratch_need()).
    bool energy_gradient_noalloc(const double *angles, double *E, double *deriv, std::string &err);
    static size_t noalloc_scratch_doubles();
    size_t noalloc_scratch_need() const;
    // psi = U(angles) psi_0, overlap = <proj|psi> and, by the adjoint sweep
    // with lambda = proj, deriv[k] = d<proj|psi>/d(angle_k) per operator
    // (FusedEvolve::evolveDerivativeProj, src/fusedevolve.cpp:2814-2889);
    // proj: N amplitudes in sector order.
    bool overlap_gradient(const double *angles, const double *proj, double *overlap, double *deriv,
                          std::string &err);
    // Per-parameter Hessian (row-major P x P), gradient and energy; metric
    // (optional, P x P): g[mu][nu] = <dpsi/dtheta_mu|dpsi/dtheta_nu>.
    bool hessian(const double *angles, double *hess, double *grad, double *E, std::string &err,
                 double *metric = nullptr);
    // The same for the overlap f = <proj|psi(angles)> (proj: N amplitudes in
    // sector order): d2f/dtheta_mu dtheta_nu, df/dtheta, f
    // (FusedEvolve::evolveHessianProj, src/fusedevolve.cpp:2892-3183).
    bool hessian_proj(const double *angles, const double *proj, double *hess, double *grad, double *overlap,
                      std::string &err);
    // The current psi over the sector (its real plane for complex amplitudes).
    bool download_state(double *out, std::string &err);
    // H psi of the current psi over the sector.
    bool download_ham_state(double *out, std::string &err);
    // y = H x for one host vector over the sector.
    bool apply_ham(const double *x, double *y, std::string &err);
    // The states psi_v = U(angles_v) psi_0 of nvec per-operator angle
    // vectors (row-major nvec x M) and their Gram matrices (row-major
    // nvec x nvec): Hmat[v][w] = <psi_v|H|psi_w>, Smat[v][w] = <psi_v|psi_w>.
    bool subspace(int nvec, const double *angles, double *Hmat, double *Smat, std::string &err);
    // Batched: nvec angle vectors (row-major nvec x M).  deriv may be null.
    bool batch(int nvec, const double *angles, double *E, double *deriv, std::string &err);
    // The same in three steps, so that repeated evaluations run with their
    // inputs resident on the device: upload (sin/cos tables of the angles),
    // run (enqueue only; no host synchronisation), download (synchronise,
    // copy energies and per-operator derivatives).
    bool batch_upload(int nvec, const double *angles, std::string &err);
    bool batch_run(bool with_gradient, std::string &err);
    // The batch sequence enqueued on the engine's stream without a graph
    // (batch_run replays it from a hipGraph); for callers that capture several
    // engines into one graph.
    bool batch_enqueue(bool with_gradient, std::string &err);
    // The engine's HIP stream (hipStream_t).
    void *stream() const;
    bool batch_download(double *E, double *deriv, std::string &err);
    int batch_size() const;
    // Times every rocBLAS solution of the lane-layout sigma's two same-spin
    // DGEMMs on the uploaded batch's buffers (contents are overwritten) and
    // keeps the fastest of those whose output on a fixed input agrees with
    // the default solution's (1e-10 relative); report: JSON with the counts,
    // rejections and times.  Does nothing unless the batch runs the
    // lane-layout matrix-free sigma.
    bool tune_sigma_gemms(std::string &report, std::string &err);
    bool synchronize(std::string &err);

    // HIP-event timing of kernel classes on the engine's stream:
    // 0 Hamiltonian application (sparse part), 1 forward sweep, 2 adjoint
    // sweep, 3 dense same-spin GEMMs of the lane-layout sigma.  Lane-layout
    // sweeps record one interval per launch.
    void set_timing(bool on);
    bool timing_on() const;
    // Changes whenever previously captured launch sequences are stale
    // (new operators, Hamiltonian, start state or engine options).
    int generation() const;
    bool timing(int cls, double *ms, long long *launches, std::string &err);

    bool set_option(int key, int value, std::string &err);
    int option(int key) const;

    // Diagnostics for the benchmark and tests.
    long long ham_nnz() const;
    int ham_kind() const; // 0 none, 1 rows, 2 uniform-width rows, 3 matrix-free
    int num_segments() const;
    int num_generic_ops() const;

private:
    bool second_derivatives(const double *angles, const double *proj, double *hess, double *grad, double *val,
                            double *metric, std::string &err);
    std::unique_ptr<EngineImpl> d_;
};

} // namespace ae
