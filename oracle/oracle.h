/* CPU oracle for the AnsatzEvolve statevector path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference algorithm (apu727/AnsatzEvolve,
 * src/fusedevolve.cpp, src/hamiltonianmatrix.cpp, src/operatorpool.{h,cpp},
 * src/sparsematrix.h, src/tupsquantities.cpp).  Only the test-suite, the smoke
 * check and the benchmark's CPU-baseline leg may load it, and only as the
 * checker.  The product library never links it.
 *
 * Bit layout: bit q of a basis state is the occupation of spin-orbital q; bits
 * 0..n-1 are spin-down (beta), bits n..2n-1 spin-up (alpha)
 * (src/operatorpool.cpp:221-222).
 */
#ifndef ANSATZ_ORACLE_H
#define ANSATZ_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_SECTOR_FULL = 0, OR_SECTOR_NUMBER = 1, OR_SECTOR_SZ = 2 };

/* Opaque engine: sector + Hamiltonian + operators + parameter map + start. */
typedef struct or_engine or_engine;

/* Sector enumeration (ascending full-space index). Returns N, or -1. */
long or_sector_size(int nq, int mode, int nel, int nup, int ndown);
long or_sector_states(int nq, int mode, int nel, int nup, int ndown, uint32_t *out);

/* applyExcToBasisState_ (real mode).  Returns +1 / -1 and the partner in
 * *out; *out == state means the operator does not act on this state.
 * Returns 0 for an operator the real build rejects (number operators). */
int or_apply_exc(uint32_t state, const int8_t a[4], uint32_t *out);

/* Engine construction.
 *   ops: nops x 4 zero-based spin-orbital indices, -1 = absent.
 *   order: zero-based parameter per operator; scale: per operator (NULL = 1).
 *   start_idx/start_coeff: zero-based full-space basis states of psi_0.
 * Returns NULL on error.  The sector mode follows the reference's symmetry
 * detection (src/AnsatzManager.cpp:88-100, 202-237, 271-374). */
or_engine *or_engine_create(int nq, int nstart, const int *start_idx, const double *start_coeff,
                            int nops, const int *ops, const int *order, const double *scale);
void or_engine_free(or_engine *e);
int or_engine_mode(const or_engine *e);
long or_engine_dim(const or_engine *e);
int or_engine_nparams(const or_engine *e);
const uint32_t *or_engine_states(const or_engine *e);

/* Hamiltonian from zero-based full-space triplets
 * (src/hamiltonianmatrix.cpp:628-659).  Duplicates are summed. */
int or_engine_set_triplets(or_engine *e, long ntrip, const int *i0, const int *j0, const double *v);
/* Hamiltonian from integrals h[n*n], g[n^4] in the reference file layout
 * (src/hamiltonianmatrix.cpp:402-554).  explicit_matrix=1 builds the sector
 * matrix (constructFromSecQuant, :556-601); 0 keeps the operator list and
 * applies it matrix-free (:915-1009). */
int or_engine_set_integrals(or_engine *e, int norb, const double *h, const double *g, int explicit_matrix);
long or_engine_nnz(const or_engine *e);

/* y = M^T x over the sector (src/hamiltonianmatrix.cpp:734, :884, :950). */
int or_ham_apply(const or_engine *e, const double *x, double *y);

/* angles: per-operator (length nops) */
int or_evolve(const or_engine *e, const double *angles, double *psi);
int or_energy(const or_engine *e, const double *angles, double *E);
int or_energy_gradient(const or_engine *e, const double *angles, double *E, double *grad_comp);
int or_hessian(const or_engine *e, const double *angles, double *hess_colmajor, double *grad_comp, double *E);
/* overlap f = proj . psi over the sector (proj: N amplitudes), its gradient
 * and Hessian (column-major) */
int or_overlap_gradient(const or_engine *e, const double *angles, const double *proj, double *ov, double *grad_comp);
int or_hessian_proj(const or_engine *e, const double *angles, const double *proj, double *hess_colmajor,
                    double *grad_comp, double *ov);

/* Complex mode (the reference's useComplex build): interleaved amplitudes
 * (2i = Re, 2i + 1 = Im).  or_apply_exc_complex works on doubled states
 * (bits << 1) | part.  A complex start replaces the real one. */
int or_apply_exc_complex(uint32_t state2, const int8_t a[4], uint32_t *out);
int or_engine_set_start_complex(or_engine *e, int nstart, const int *idx, const double *re, const double *im);
int or_evolve_complex(const or_engine *e, const double *angles, double *psi_interleaved);
int or_energy_gradient_complex(const or_engine *e, const double *angles, double *E, double *grad_comp);

/* parameter map D (src/tupsquantities.cpp:643-667): angles_ops = D theta */
int or_decompress_angles(const or_engine *e, const double *theta, double *angles_ops);

#ifdef __cplusplus
}
#endif
#endif
