/* AnsatzEvolve statevector engine for MI355X (gfx950) -- C ABI.
 *
 * Part 1 is a drop-in for the reference's C interface
 * (apu727/AnsatzEvolve src/Generated/AnsatzSynthInterface.h:14-26, generated
 * from src/AnsatzSynthInterface.f90 by gfortran -fc-prototypes): same names,
 * same signatures, same argument meaning, same return codes.  Existing C and
 * Fortran callers (the bind(C) module in src/AnsatzSynthInterface.f90) link
 * libansatzsynth.so instead of libAnsatzSynthInterface.a +
 * libcppAnsatzSynthLib.a and need no source change.
 *
 * Part 2 adds entry points the reference's C interface lacks (integral
 * Hamiltonians, nuclear energy, fused and batched evaluations).  They follow
 * the same conventions: opaque ctx from init(), caller-owned host arrays that
 * are copied, int status codes, no exceptions, one lock per ctx.
 *
 * Conventions (identical to the reference):
 *  - basis index = bit string + 1 (one-based); bit q = spin-orbital q; bits
 *    0..n-1 spin-down, n..2n-1 spin-up;
 *  - operators[nparams*j + i] = one-based component j of operator i (0 = absent),
 *    (a,b,c,d) = a^+_a a^+_b a_c a_d, single (a,b,0,0) = a^+_a a_b;
 *  - angles may be given per operator (length = number of operators) or per
 *    parameter (length = max(orderfile)) wherever the reference accepts both;
 *  - gradients and Hessians are per parameter; Hessians are column-major,
 *    hessian[nangles*j + i] = d2E / dtheta_i dtheta_j;
 *  - energies exclude nuclear repulsion unless setNuclearEnergy was called.
 */
#ifndef ANSATZSYNTH_H
#define ANSATZSYNTH_H

#include <stddef.h>
#ifdef __cplusplus
#include <complex>
#define ANSATZ_DOUBLE_COMPLEX std::complex<double>
extern "C" {
#else
#define ANSATZ_DOUBLE_COMPLEX double _Complex
#endif

/* ---------------- Part 1: the reference C interface ---------------- */

/* ctx is the address of the handle (void**), as Fortran passes c_ptr by
 * reference; sets *ctx = NULL.  0 ok (also when *ctx is already NULL),
 * 1 ctx == NULL.  Replaces src/AnsatzSynthInterface.cpp:42-58. */
int cleanup(void *ctx);

/* 1 ctx NULL, 2 angles NULL, 3 nangles < 1 or evaluation failed.
 * Replaces src/AnsatzSynthInterface.cpp:313-344. */
int getEnergy(int nangles, const double *angles, double *energy, void *ctx);

/* Dense full-Fock final state (index = bit string).  1 ctx NULL, 2 angles
 * NULL, 3 finalstate NULL, 4 nbasisvectors < 1, 5 nangles < 1, 6 failure,
 * 7 nbasisvectors != 2^numqubits.  Replaces src/AnsatzSynthInterface.cpp:362-420. */
int getFinalState(int nangles, const double *angles, int nbasisvectors, double *finalstate, void *ctx);

/* The final state over the full space as complex amplitudes (imaginary
 * parts zero for a real context).  Codes as getFinalState.  Replaces the
 * complex build's src/AnsatzSynthInterface.cpp:439-486 (its real build
 * returns 4; this library has one build, and 4 means nbasisvectors < 1). */
int getFinalStateComplex(int nangles, const double *angles, int nbasisvectors,
                         ANSATZ_DOUBLE_COMPLEX *finalstate, void *ctx);

/* Per-parameter gradient.  1 ctx NULL, 2 angles NULL, 3 gradient NULL,
 * 4 nangles < 1, 5 failure, -6 nangles != number of parameters.
 * Replaces src/AnsatzSynthInterface.cpp:519-562. */
int getGradient_COMP(int nangles, const double *angles, double *gradient, void *ctx);

/* Per-parameter Hessian, column-major.  Codes as getGradient_COMP.
 * Replaces src/AnsatzSynthInterface.cpp:578-629. */
int getHessian_COMP(int nangles, const double *angles, double *hessian, void *ctx);

/* New context; free with cleanup(&ctx).  Replaces src/AnsatzSynthInterface.cpp:28-33. */
void *init(void);

/* = setExcitationScale(..., NULL, ctx).  Replaces src/AnsatzSynthInterface.cpp:64-67. */
int setExcitation(int nparams, const int *operators, const int *orderfile, void *ctx);

/* 1 ctx NULL, 2 nparams < 1, 3 operator index - 1 > 127, 4 operators not
 * stored (e.g. after the first evaluation), 5 orderfile entry > nparams,
 * 6 parameter map not stored.  Replaces src/AnsatzSynthInterface.cpp:84-140. */
int setExcitationScale(int nparams, const int *operators, const int *orderfile, const double *scalefactor,
                       void *ctx);

/* One-based full-Fock triplets (i, j, v), duplicates summed; the operator is
 * applied as y = M^T x.  1 ctx NULL, 2 n < 1, 3 not stored.
 * Replaces src/AnsatzSynthInterface.cpp:154-187. */
int setHamiltonian(int n, const int *iindexes, const int *jindexes, const double *coeffs, void *ctx);

/* One-based basis indices.  1 ctx NULL, 2 n < 1 or numqubits < 1,
 * 3 not stored.  Replaces src/AnsatzSynthInterface.cpp:201-236. */
int setInitialState(int numqubits, int n, const int *iindexes, const double *coeffs, void *ctx);

/* Complex start state: 1 ctx NULL, 2 n < 1, numqubits < 1 or a NULL
 * array (the reference does not check the arrays), 3 rejected
 * (as setInitialState).  The context then evaluates complex amplitudes, as
 * the reference's complex build does: number operators (a^+_p a_p,
 * a^+_p a^+_q a_p a_q) are allowed and multiply the states where they act
 * by exp(+i angle); getFinalState returns the real part.  Energy and
 * gradient entry points work as for real amplitudes; the Hessians, overlap
 * gradients, batches, optimisers, subspace and property entry points and
 * getEnergyGradientNoAlloc return their failure code.  A context whose
 * operators include number operators is complex even with a real start.
 * Replaces the complex build's src/AnsatzSynthInterface.cpp:251-285. */
int setInitialStateComplex(int numqubits, int n, const int *iindexes, const ANSATZ_DOUBLE_COMPLEX *coeffs,
                           void *ctx);

/* Global: non-zero logs every later call's arguments to stderr.
 * Replaces src/AnsatzSynthInterface.cpp:18-21. */
void setTraceInterfaceCalls(int val);

/* ---------------- Part 2: extensions ---------------- */

/* Molecular-orbital integrals instead of triplets: h[norb*norb] row-major,
 * g[norb^4] in the layout of the reference's _twoEInts.bin
 * (src/hamiltonianmatrix.cpp:402-554; the reference reads them from files
 * via storeRunPath, src/AnsatzManager.h:177).  numqubits must be 2*norb.
 * 1 ctx NULL, 2 norb < 1 or h/g NULL, 3 not stored (after construction). */
int setHamiltonianIntegrals(int norb, const double *h, const double *g, void *ctx);

/* Nuclear repulsion added to every energy (stateAnsatzManager::storeNuclearEnergy,
 * src/AnsatzManager.cpp:400-413).  1 ctx NULL, 3 not stored. */
int setNuclearEnergy(double energy, void *ctx);

/* How an integral Hamiltonian is applied: 0 automatic (matrix-free from 12
 * orbitals or when the explicit rows do not fit in device memory, explicit
 * rows otherwise), 1 explicit, 2 matrix-free.
 * 1 ctx NULL, 2 bad mode, 3 after construction. */
int setHamiltonianStorage(int mode, void *ctx);

/* One evaluation of energy and per-parameter gradient at the same angles
 * (FusedEvolve::evolve + evolveDerivative(&E), src/fusedevolve.cpp:2105-2159,
 * 2263-2346).  Codes as getGradient_COMP; energy may be NULL. */
int getEnergyGradient(int nangles, const double *angles, double *energy, double *gradient, void *ctx);

/* getEnergyGradient for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  The call makes no device allocation, no
 * graph and no library call: the angle tables, the partial sums and the
 * results live in one fixed device array of 32 MiB (4 Mi doubles) that each
 * GPU holds until the process ends; the states use the context's own
 * buffers, made when the context was built, so the context must have been
 * evaluated once by an ordinary call.  Every call on a GPU shares that
 * GPU's array: calls from this process are serialised, calls that overlap
 * on one GPU from elsewhere race.  The Hamiltonian is applied by the
 * explicit rows or the all-in-one matrix-free kernel.  angles per
 * parameter.  Codes as getGradient_COMP, and 6 the context is not built
 * yet, 7 the set-up needs more than the array holds (parameter count times
 * the sweep's partial rows, about 455 x 4096 at most). */
int getEnergyGradientNoAlloc(int nangles, const double *angles, double *energy, double *gradient, void *ctx);

/* Batched energies for nvec angle vectors (row-major nvec x nangles), as the
 * reference's stateAnsatzManager::getExpectationValues
 * (src/AnsatzManager.cpp:709-754).  1 ctx NULL, 2 angles/energies NULL,
 * 3 nvec < 1 or nangles < 1, 5 failure. */
int getEnergies(int nvec, int nangles, const double *angles, double *energies, void *ctx);

/* Batched energy + gradient, nvec vectors; gradients row-major nvec x nangles.
 * Codes as getEnergies, -6 nangles != number of parameters. */
int getEnergyGradients(int nvec, int nangles, const double *angles, double *energies, double *gradients,
                       void *ctx);

/* Device-resident batch, for callers that evaluate the same angle vectors
 * repeatedly (benchmarks, line searches re-run after a set-up change) or
 * want to overlap host work with the device:
 *   uploadAngleBatch  copies nvec angle vectors to the device (sin/cos tables);
 *   runAngleBatch     enqueues energies (and gradients if withgradient != 0)
 *                     for the uploaded batch and returns without waiting;
 *   downloadAngleBatch waits and copies energies / per-parameter gradients
 *                     (gradients may be NULL).
 * Codes: 1 ctx NULL, 2 pointer NULL, 3 nvec or nangles < 1, 5 failure,
 * -6 gradients requested with nangles != number of parameters. */
int uploadAngleBatch(int nvec, int nangles, const double *angles, void *ctx);
int runAngleBatch(int withgradient, void *ctx);
int downloadAngleBatch(double *energies, double *gradients, void *ctx);

/* Kernel timing with HIP events on the ctx's stream.  setKernelTiming(1)
 * resets and starts recording; getKernelTiming waits for the stream and
 * returns the summed milliseconds and launch count of one kernel class:
 * 0 Hamiltonian application (its sparse kernel), 1 forward sweep,
 * 2 adjoint sweep, 3 dense same-spin GEMMs of the batched matrix-free
 * sigma.  Batched (lane-layout) sweeps count every kernel launch, the
 * other paths one interval per sweep.  1 ctx NULL, 2 bad class, 5 failure. */
int setKernelTiming(int on, void *ctx);
int getKernelTiming(int kernelclass, double *milliseconds, long long *launches, void *ctx);

/* The plain optimiser (stateAnsatzManager::optimise, src/AnsatzManager.cpp:756-773,
 * which is C++/Python-only in the reference): Newton steps with an
 * |eigenvalue|^-1 inverse Hessian, a cubic-model line search and
 * negative-curvature escape steps, until |gradient| < 1e-12 or no further
 * progress (src/tupsquantities.cpp:704-1236, 1467-1598, 1940-2011).
 * angles (length = operators or parameters) are the start and receive the
 * result in the same form; energy receives the final energy.  When
 * stopenergy is not NaN the run also stops once an energy <= stopenergy is
 * seen.  If trace != NULL, up to maxtrace rows of (seconds, energy,
 * |gradient|, Hessian evaluations, energy evaluations) are written and
 * *ntrace is set.  1 ctx NULL, 2 pointer NULL, 3 nangles < 1, 5 failure. */
int optimise(int nangles, double *angles, double *energy, double stopenergy, int maxtrace, double *trace,
             int *ntrace, void *ctx);

/* The reference's iterative optimiser (TUPSQuantities::iterativeTups with
 * runNewtonMethodProjected, src/tupsquantities.cpp:2044-2145, 1238-1465;
 * the CLI's iterativeoptimise mode, C++-only in the reference): outer
 * iterations freeze the projection p = H psi(previous angles) and minimise
 * <p|psi(theta)> by Newton steps (|eigenvalue|^-1 inverse Hessian of the
 * overlap, cubic-model line search with forward steps), until the projected
 * gradient's norm is below 1e-5.  ediis != 0 extrapolates the outer
 * iterates by energy-DIIS over the last 15 (off in the reference).  angles,
 * energy, trace as optimise (one row per outer iteration).
 * 1 ctx NULL, 2 pointer NULL, 3 nangles < 1, 5 failure. */
int iterativeOptimise(int nangles, double *angles, double *energy, int ediis, int maxtrace, double *trace,
                      int *ntrace, void *ctx);

/* The Hamiltonian in the span of the states of nsets angle vectors
 * (TUPSQuantities::doSubspaceDiagonalisation, src/tupsquantities.cpp:
 * 1600-1768; the CLI's subspacediag mode): H_ij = <psi_i|H|psi_j> and
 * S_ij = <psi_i|psi_j>, Loewdin-symmetrised S^-1/2 H S^-1/2 with the
 * eigenvalues of S below 1e-14 dropped (those directions give zero
 * eigenvalues).  angles: nsets rows of nangles (per operator or per
 * parameter).  eigenvalues: nsets, ascending; hmat, smat (nsets x nsets)
 * and nullrank (the dropped directions) may be NULL.
 * 1 ctx NULL, 2 angles or eigenvalues NULL, 3 nsets or nangles < 1,
 * 5 failure. */
int subspaceDiagonalise(int nsets, int nangles, const double *angles, double *eigenvalues, double *hmat,
                        double *smat, int *nullrank, void *ctx);

/* Properties of the states of nsets angle vectors
 * (TUPSQuantities::writeProperties, src/tupsquantities.cpp:185-568; the
 * CLI's writeproperties mode).  props: nsets rows of 12 + 2 noverlaps
 * values: 0 energy (electronic), 1 energy + nuclear energy, 2..4 Hessian
 * eigenvalues below -1e-10 / within 1e-10 of zero / above 1e-10, 5..7 the
 * same for the Hessian's diagonal, 8..9 eigenvalues of the metric
 * g_mu,nu = <dpsi/dtheta_mu|dpsi/dtheta_nu> above 1e-10 / within 1e-10,
 * 10 metric diagonal entries within 1e-10 of zero, 11 |gradient|, then per
 * eigenvector t of H (lowest first) the overlap <v_t|psi> and its magnitude.
 * nohess != 0 skips the Hessian and metric (columns 2..10 stay 0).
 * hessians, metrics (nsets x P x P, may be NULL): the matrices; eigenvalues
 * (noverlaps): the lowest eigenvalues of H (Lanczos with the Hamiltonian on
 * the device).  1 ctx NULL, 2 pointer NULL, 3 nsets or nangles < 1 or
 * noverlaps outside 0..16, 5 failure. */
int getProperties(int nsets, int nangles, const double *angles, int noverlaps, int nohess, double *props,
                  double *hessians, double *metrics, double *eigenvalues, void *ctx);

/* Gradient per operator (the decompressed form of the C++/Python
 * stateAnsatzManager::getGradient, src/AnsatzManager.h:221): dE/dx_k for
 * the angle x_k of operator k; gradient receives one value per operator
 * (getDimensions).  angles: per operator or per parameter.  Codes as
 * getGradient_COMP (1 ctx, 2 angles, 3 gradient NULL, 4 nangles < 1,
 * 5 failure). */
int getGradientOperators(int nangles, const double *angles, double *gradient, void *ctx);

/* Overlap <proj|psi(angles)> and its per-parameter gradient, by the adjoint
 * sweep with the projection vector in place of H psi
 * (FusedEvolve::evolveDerivativeProj, src/fusedevolve.cpp:2814-2889; the
 * reference's projected optimiser uses it, src/tupsquantities.cpp:2075-2095).
 * projvector: dense over the full space, nbasisvectors = 2^numqubits (as
 * getFinalState); overlap may be NULL.  angles per parameter.
 * 1 ctx NULL, 2 angles or projvector NULL, 3 gradient NULL, 4 nangles or
 * nbasisvectors < 1, 5 failure (also nbasisvectors != 2^numqubits),
 * -6 nangles != number of parameters. */
int getOverlapGradient(int nangles, const double *angles, int nbasisvectors, const double *projvector,
                       double *overlap, double *gradient, void *ctx);

/* Hessian of the overlap f = <proj|psi(angles)>: hessian[nangles*j+i] =
 * d2f/dtheta_i dtheta_j (column-major, as getHessian_COMP), its gradient and
 * f itself (FusedEvolve::evolveHessianProj, src/fusedevolve.cpp:2892-3183;
 * the Hessian of the reference's projected Newton step,
 * src/tupsquantities.cpp:1238-1465).  projvector as getOverlapGradient;
 * gradient and overlap may be NULL.  Codes as getOverlapGradient, with
 * 3 = hessian NULL. */
int getHessianProj(int nangles, const double *angles, int nbasisvectors, const double *projvector,
                   double *hessian, double *gradient, double *overlap, void *ctx);

/* Sector dimension, number of operators and parameters after construction
 * (constructs if needed).  Any pointer may be NULL.  1 ctx NULL, 5 failure. */
int getDimensions(long long *sectordim, int *noperators, int *nparams, void *ctx);

/* Execution choices of the device engine; none changes a result beyond the
 * order of floating-point sums.  Keys:
 *   0 sweep kind: 0 register orbits (+ LDS tiles), 1 LDS-staged orbits,
 *     2 LDS tiles only;
 *   1 lane-layout threshold: batches of >= value vectors sweep with one
 *     vector per wavefront lane (0: never);
 *   2 hipGraph replay: 1 on, 0 eager launches;
 *   3 fusion depth: most operators per pass (0: the kernels' own limit);
 *   4 LDS budget in bytes of a tile segment (0: the device's limit);
 *   5 (read only) Hamiltonian form: 1 rows, 2 uniform-width rows,
 *     3 matrix-free;
 *   6 (read only) fused segments per sweep (vector-major sweeps);
 *   7 lane-layout adjoint sweep: disjoint 2-bit orbital groups per spin
 *     fused into one pass (1, 2 or 3, default 1);
 *   8 (read only) fused segments per lane-layout adjoint sweep;
 *   9 the same as 7 for the lane-layout forward sweep;
 *  10 stored forward states for the lane-layout gradient: 1 (default) keeps
 *     the state after every segment (out-of-place forward sweep, adjoint
 *     sweep reading them instead of writing psi back) when both sweeps share
 *     their segments and the states fit in half the free device memory;
 *     0 in-place sweeps;
 *  11 (read only) 1 when the uploaded batch runs with stored states;
 *  12 lane-layout matrix-free sigma: 1 the variant with the source rows
 *     split in two double-buffered halves when it applies, 0 (default) the
 *     one-row variant;
 *  13 (read only) 1 when the two-half variant applies (12 or fewer
 *     orbitals, as many beta electrons as empty beta orbitals);
 *  14 lane-layout matrix-free sigma: 1 (default) the variant whose 16-lane
 *     rows each hold all targets with one occupation of the upper six beta
 *     orbitals, when it applies (it then takes precedence over key 12),
 *     0 the variants of key 12;
 *  15 (read only) 1 when the grouped variant applies (12 orbitals, 6 beta
 *     electrons);
 *  16, 17 rocBLAS solution index of the lane-layout sigma's beta-part and
 *     alpha-part DGEMMs (0, the default: the library's own choice; the
 *     autotuner times every solution and keeps the fastest; an index the
 *     library does not offer falls back to 0);
 *  18 1 (default): the lane-layout forward sweep acts only on the orbits
 *     that the start state and the operators before each segment can reach
 *     (all other amplitudes are exact zeros), and the adjoint sweep reads
 *     stored forward states only there; 0: every orbit.  Results are the
 *     same bit for bit;
 *  19, 20, 21 (read only, after a lane-layout batch upload) amplitudes per
 *     vector and step, summed over the segments: read and written by the
 *     forward sweep (19), of the adjoint vector read and written by the
 *     adjoint sweep (20), of the stored forward states read by it (21).
 *  25 window passes for the lane-layout sweeps: w = 2 to 8 runs each sweep
 *     as passes over windows of w consecutive spatial orbitals, each pass
 *     staging the determinants that agree outside its window in LDS and
 *     applying every operator run inside the window there (one read and
 *     write of the state per pass instead of per run); 0: orbit segments;
 *     -1 (default): w = 4 when the tiles fit in registers (key 28), else
 *     orbit segments.  Used when every operator falls into a run on at most
 *     two alpha and two beta orbitals spanning at most w orbitals;
 *  26 waves per block of the window-pass kernels: 4 (default) or 8;
 *  27 (read only) window passes per sweep, 0 when they are not in use;
 *  28 (read only) 1 when the window passes keep each tile in registers
 *     (width 4 and every operator run on one adjacent orbital pair in both
 *     spins), 2 when in addition every run is a TUPS unit of five operators
 *     (beta single, alpha single, paired double, beta single, alpha single:
 *     straight-line phases), 0 when they stage tiles in LDS or are not used;
 *  29 window passes for the vector-major sweeps (single vectors and
 *     batches below key 1): -1 (default) from 2^18 determinants on, 0
 *     never, 1 at any size.  They use the register tiles of key 28 (1 or 2)
 *     with one tile per lane, applied in place; key 24 then counts them;
 *  30 1 (default): a batch evaluated with its gradient on register window
 *     tiles takes the energies <psi|H psi> in the first adjoint pass (whose
 *     tiles hold the final state and H psi) instead of a separate dot over
 *     both vectors; 0: the separate dot;
 *  31 1 (default): the Hessian's tangent sweep (getHessian_COMP,
 *     getHessianProj) runs as the window passes when they hold TUPS units
 *     in registers (key 28 = 2); 0: the two-group segments;
 *  32 1 (default): the vector-major window passes (key 29) copy each alpha
 *     class's rows of the vector to LDS with coalesced loads and take the
 *     tiles from there, where six rows (twice for the adjoint) fit in LDS;
 *     0: each tile straight from memory.  Reads back 1 when in use, 2 when
 *     requested but the rows do not fit, 0 when off;
 *  33 -1 (default): a lane-layout batch whose states are symmetric under the
 *     exchange of the alpha and beta strings takes the beta-beta part of the
 *     grouped sigma (12 orbitals) from the alpha-alpha DGEMM's result read
 *     transposed instead of a second DGEMM.  It applies when both spins have
 *     the same strings and the same same-spin matrix (bit for bit), the
 *     start vector satisfies x[a][b] == x[b][a], and every operator is its
 *     own mirror image under the exchange or sits next to its mirror (acting
 *     on the other spin only) with the same angle in every vector of the
 *     batch, as in TUPS units; checked at each upload.  Single-vector calls
 *     (getEnergyGradient) and vector-major batches with such angles compute
 *     one triangle of the vector-major sigma's targets and copy the rest;
 *     the Hessian's H B does so too when mirror operators also share their
 *     parameter.  0: never (takes
 *     effect at once; -1 from the next upload).  Reads back 1 when in use for
 *     the uploaded batch, 2 when allowed but not applicable, 0 when off.
 * Changing keys 0, 3, 4, 7, 9, 10, 18, 25 or 26 re-segments the operators or changes the
 * buffers; an uploaded angle batch must then be uploaded again.  1 ctx NULL, 2 pointer NULL,
 * 5 rejected value or failure. */
int setEngineOption(int key, int value, void *ctx);
int getEngineOption(int key, int *value, void *ctx);

/* Autotuner of the batched E+grad sequence for nvec angle vectors
 * (nangles = operators or parameters per vector): runs every candidate
 * (sweep kind x fusion depth x batch layout), reps timed replays each,
 * checks that all candidates give the same energies and derivatives, and
 * leaves the fastest applied.  If cachefile is not NULL, a recorded choice
 * for this shape (sector dimension, operators, Hamiltonian form, nvec) is
 * applied without measuring, and a new choice is appended as one JSON line.
 * report (if not NULL) receives a JSON summary, truncated to reportlen - 1
 * characters.  1 ctx NULL, 2 angles NULL, 3 nvec < 1, 5 failure. */
int autotune(int nvec, int nangles, const double *angles, const char *cachefile, int reps, char *report,
             int reportlen, void *ctx);

/* Several contexts evaluated by one launch: each context (its own
 * Hamiltonian, sector size and operators; constructed) gets an angle batch
 * with multiBatchUpload, multiBatchRun enqueues all of them as parallel
 * branches of one hipGraph (captured once per shape) and returns without
 * waiting, multiBatchDownload waits and copies one context's energies and
 * per-parameter gradients (may be NULL).  A context may belong to one
 * multi-batch at a time; the contexts stay owned by the caller and must
 * outlive the multi-batch.  No reference counterpart (the reference
 * evaluates one Hamiltonian per manager).
 * Codes: 1 NULL handle, 2 bad index or NULL pointer, 3 nvec or nangles < 1,
 * 5 failure, -6 gradients requested with nangles != number of parameters. */
void *multiBatchCreate(int nctx, void *const *ctxs);
int multiBatchUpload(void *mb, int index, int nvec, int nangles, const double *angles);
int multiBatchRun(void *mb, int withgradient);
int multiBatchDownload(void *mb, int index, double *energies, double *gradients);
int multiBatchDestroy(void *mb);

/* Run files.  The reference reads them in its command-line program
 * (src/main.cpp:233-380 with src/TUPSLoadingUtils.cpp) and, for the
 * Hamiltonian only, through stateAnsatzManager::storeRunPath
 * (src/AnsatzManager.cpp:447-466, 185-199), which has no C-ABI binding.
 *
 * setRunPath: with no Hamiltonian set, construction loads it from
 * <runpath>_Ham_Coeff.dat + _Ham_Index.dat (text triplets, one-based
 * full-space indices) when both exist, else from _oneEInts.bin +
 * _twoEInts.bin (FP64 integrals, n^2 and n^4, n = qubits / 2), as the
 * reference's file constructor (src/hamiltonianmatrix.cpp:660-705).
 * 1 ctx NULL, 2 runpath NULL or empty, 3 already constructed. */
int setRunPath(const char *runpath, void *ctx);
/* The whole set-up from <prefix>_Initial.dat (bitstring, coefficient),
 * _Operators.dat (four one-based indices per operator), _Order.dat
 * ("parameter[,scale]" per operator), the Hamiltonian as in setRunPath,
 * and (withnuclear != 0) _Nuclear_Energy.dat, added to every energy (a
 * missing file counts as 0, src/TUPSLoadingUtils.cpp:156-174).
 * 1 ctx NULL, 2 prefix NULL or empty, 3 a file missing or malformed (or
 * operator and order counts differ), 4 a set-up step failed (e.g. already
 * constructed).  The Hamiltonian files are read at construction. */
int loadRunFiles(const char *prefix, int withnuclear, void *ctx);
/* The parameter sets of a _Parameters.dat file (a count line, an "Energy of
 * minimum ... = E" line, then the values; src/TUPSLoadingUtils.cpp:176-234).
 * Sets whose value count differs from their count line are skipped.  Sets
 * *nsets and *nparams (the longest set); with maxsets > 0 fills energies[s]
 * and params[s * maxparams + i] (zero padded).  No ctx.  2 NULL path or
 * count pointer, 3 file missing or malformed, 6 more sets or parameters
 * than the buffers hold (counts still set) or NULL buffers. */
int readParameterSets(const char *path, int maxsets, int maxparams, int *nsets, int *nparams, double *energies,
                      double *params);

/* ---------------- Part 3: learned warm start (no reference counterpart) ----
 * A 3-layer FP32 network x -> tanh(x W0 + b0) -> tanh(. W1 + b1) -> . W2 + b2
 * proposing initial angles from integral features, with device kernels for
 * inference and training.  Parameters are laid out W0[in][hid], b0, W1, b1,
 * W2[hid][out], b2; arrays are row-major [batch][features].  Loss is the
 * masked mean squared error sum m (y - t)^2 / sum m.
 * Codes: 1 NULL handle/pointer, 3 bad size, 5 device failure. */
void *mlpCreate(int nin, int nhidden, int nout, unsigned long long seed);
int mlpDestroy(void *mlp);
long long mlpParameterCount(void *mlp);
int mlpGetParameters(void *mlp, float *params);
int mlpSetParameters(void *mlp, const float *params);
int mlpForward(void *mlp, int batch, const float *x, float *y);
int mlpGradient(void *mlp, int batch, const float *x, const float *target, const float *mask, float *loss,
                float *grads);
int mlpTrainStep(void *mlp, int batch, const float *x, const float *target, const float *mask, float lr,
                 float *loss);

#ifdef __cplusplus
}
#endif
#endif /* ANSATZSYNTH_H */
