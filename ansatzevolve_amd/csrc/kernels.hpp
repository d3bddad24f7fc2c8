// HIP kernels of the statevector engine (gfx950, wave64, FP64 on VALU).
//
// Data layout in HBM (S_z sector, n spatial orbitals, k_up / k_down electrons):
//   sector index i = rank_up(alpha) * N_down + rank_down(beta), i.e. the
//   reference's ascending full-index order (src/operatorpool.cpp:207-242),
//   alpha-major; a state vector is N = N_up * N_down doubles, each alpha
//   string owns one contiguous row of N_down amplitudes.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.hpp"

namespace ae {

constexpr int kBlock = 256;
constexpr int kMaxTileRows = 64;

// ---------------------------------------------------------------- helpers
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// Deterministic block sum (fixed tree); result valid in thread 0.  `red`
// holds kBlock/64 doubles; ends with a barrier so `red` can be reused.
__device__ __forceinline__ double block_sum(double v, double *red)
{
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) red[w] = v;
    __syncthreads();
    double t = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < kBlock / 64; i++) t += red[i];
    __syncthreads();
    return t;
}

// Pair rule of one operator on bit string b (src/fusedevolve.cpp:27-145).
// Returns 0 if the operator does not act on b, else the phase s = +-1 and
// the partner in b2.
__device__ __forceinline__ int pair_sign(const DevOp &o, uint32_t b, uint32_t &b2)
{
    const uint32_t m = b & (o.cre | o.ann);
    int s;
    if (m == o.ann) s = 1;
    else if (m == o.cre) s = -1;
    else return 0;
    b2 = b ^ o.cre ^ o.ann;
    if (__popc(b & o.mask) & 1) s = -s;
    return s * o.flip;
}

// Sector rank in any of the three sector kinds.
struct DevSector {
    int mode;  // SectorMode
    int n;     // nq / 2
    int nel;   // particle number (kNumber)
    int nb;    // N_down (kSz)
    uint32_t maskB;
    const uint32_t *states;  // N bit strings, ascending
    const int32_t *rkA;      // 2^n -> rank among alpha strings (-1 if wrong count)
    const int32_t *rkB;      // 2^n -> rank among beta strings
    const uint32_t *binom;   // 33 x 33, binom[p*33+t] = C(p, t)

    __device__ __forceinline__ long long rank(uint32_t b) const
    {
        if (mode == kSz) {
            const int a = rkA[b >> n], c = rkB[b & maskB];
            return (a < 0 || c < 0) ? -1 : (long long)a * nb + c;
        }
        if (mode == kNumber) {
            if (__popc(b) != nel) return -1;
            // colex rank = ascending-integer rank among strings of equal popcount
            long long r = 0;
            int t = 1;
            while (b) {
                const int p = __ffs(b) - 1;
                r += binom[p * 33 + t];
                t++;
                b &= b - 1;
            }
            return r;
        }
        return (long long)b;
    }
};

// Tiles of one fused segment: alpha strings that agree outside the alpha
// bits U touched by the segment form one tile; the segment's operators map
// every (alpha, beta) of a tile to the same tile.
struct SegArgs {
    int n, nb;
    uint32_t maskB;
    const uint32_t *unA, *unB;
    const int32_t *rkA, *rkB;
    const int32_t *tileStart;  // ntiles + 1
    const int32_t *rows;       // alpha ranks, grouped by tile
    const int8_t *localRow;    // alpha rank -> row within its tile
    const DevOp *ops;
    int k0, k1;
    int maxRows;
    int reverse;  // 1: apply k1-1 .. k0 (adjoint sweep), 0: k0 .. k1-1
};

// Orbit segments: consecutive operators that together touch at most two
// alpha and two beta spin-orbitals (Ua, Ub).  Determinants that agree outside
// Ua and Ub form an orbit of at most 2 x 2 members that the segment maps into
// itself; one thread owns one orbit in registers for the whole segment.
// Alpha classes: alpha strings agreeing outside Ua (1 or 2 members), the
// same for beta.  An orbit is (alpha class, beta class).
constexpr int kOrbMaxOps = 8;
struct OrbArgs {
    int n, nb;
    uint32_t maskB;
    const uint32_t *unA, *unB;
    const int32_t *aStart, *aRows;  // alpha classes: members aRows[aStart[c] ..]
    const int32_t *bStart, *bRows;  // beta classes
    // packed class descriptors {rank0, rank1 (or -1), string0, string1}
    const int4 *aDesc, *bDesc;
    int nAc, nBc;
    const DevOp *ops;
    int k0, k1;
    int reverse;
    // Orbit program of the segment (lane-layout kernel): an orbit's type is
    // (alpha shape, beta shape), shape 0 = two members, 1 = one member with
    // the segment's bits empty, 2 = one member with them full.  Word
    // prog[type * kOrbMaxOps + kk] for operator k0 + kk: bit r (r < 6) marks
    // member pair r = (0,1) (0,2) (0,3) (1,2) (1,3) (2,3) as a rotation pair,
    // bit 8 + r its phase -1 seen from the first member, phase computed with
    // the spectator bits cleared.  The spectator part is
    // (-1)^popc(spectators & mask) per operator.
    const uint32_t *prog;
    uint32_t uaLocal;   // the segment's alpha bits within an alpha string
    uint32_t ubLocal;   // its beta bits within a beta string
    uint32_t spect;     // full-string mask of the spectator bits
};

// ---------------------------------------------------------------- kernels
__global__ void k_fill(double *__restrict__ p, long long n, double v);
// Per block b: out[2b] = max |a_i - ref_i| (+inf where that is not finite),
// out[2b+1] = max |ref_i|, over a grid-stride range of n elements; kBlock threads.
__global__ void k_absdiff_max(const double *__restrict__ a, const double *__restrict__ ref, long long n,
                              double *__restrict__ out);
__global__ void k_copy(const double *__restrict__ src, double *__restrict__ dst, long long n);
// p[i] = a fixed pseudo-random value in [-1, 1) (test inputs on the device)
__global__ void k_pattern(double *__restrict__ p, long long n);

template <bool BR>
__global__ void k_orb(OrbArgs a, double *__restrict__ v0, double *__restrict__ v1, long long vstride,
                      const double2 *__restrict__ cs, long long csStride, double *__restrict__ partial, int slots,
                      long long pstride);

// The same with one workgroup per alpha class: the class's rows (1 or 2)
// are loaded into LDS with contiguous row reads, threads process beta-class
// orbits from LDS, and the rows are stored back contiguously.  Partials are
// indexed by alpha class.
template <bool BR>
__global__ void k_orbl(OrbArgs a, double *__restrict__ v0, double *__restrict__ v1, long long vstride,
                       const double2 *__restrict__ cs, long long csStride, double *__restrict__ partial, int slots,
                       long long pstride);
// dst[v*n + i] = src[i] for v < nvec
__global__ void k_bcast(const double *__restrict__ src, double *__restrict__ dst, long long n, int nvec);

// Lane layout of a batch: vectors in chunks of 64, amplitude i of vector
// 64c + l at [(c*N + i)*64 + l].  One wavefront's 64 lanes are the chunk's
// 64 vectors, so every index, class and sign computation is wave-uniform
// (scalar unit) and each amplitude access is one contiguous 512-byte load.
// Angle tables: cs[(c*M + k)*64 + l]; bracket partials:
// partial[((c*M + k)*slots + block)*64 + l], one row per block and operator
// of the segment (every block of the fixed `slots`-block grid writes its row).
constexpr int kLanes = 64;

// Amplitudes x are read from v0 and written to xout: xout == v0 in place;
// another buffer out of place (every member written, lone determinants
// copied); nullptr not written (adjoint sweep over a stored forward state).
// The same holds for k_orb2v.
template <bool BR>
__global__ void k_orbv(OrbArgs a, const double *v0, double *xout, double *__restrict__ v1, long long cstride,
                       const double2 *__restrict__ cs, int M, double *__restrict__ partial, int slots);
// Two-group orbit segments (lane layout): consecutive operators whose alpha
// bits are one of at most two disjoint 2-bit groups (or none) and likewise for
// beta -- two TUPS units on disjoint spatial pairs.  Every operator keeps the
// electron count of each group, so an orbit is the product of per-group
// classes: alpha members ma = ma1 + 2 ma2, beta members mb = mb1 + 2 mb2
// (m = 0 the group's lower bit set, 1 its upper bit; groups holding 0 or 2
// electrons have only m = 0), slot = ma + 4 mb, at most 16 members.  A
// segment is two consecutive runs: operators on the first groups, then
// operators on the second groups (two TUPS units on disjoint pairs), and
// every operator's pair pattern is the same in its four sub-orbits.
// Class descriptors: ranks[4] (-1 absent) and info = spectator bits |
// shape1 << 24 | shape2 << 26 (shape 0 two members, 1 empty, 2 full).
// Orbit type = shapeA1 + 3 shapeA2 + 9 (shapeB1 + 3 shapeB2).  An operator of
// run g (g = 0 first, 1 second) acts within the 4 sub-orbits that fix the
// other groups' members (combo c = other alpha member + 2 other beta member);
// prog[type * kOrb2MaxOps + kk] holds its 6 member-pair bits and at bits 6..11
// their phases (first member's view, spectators cleared), the same in all
// four sub-orbits.
constexpr int kOrb2MaxOps = 10;
struct Orb2Args {
    int n, nb;
    const int4 *aRanks, *bRanks;
    const uint32_t *aInfo, *bInfo;
    int nAc, nBc;
    // classes with two members in some group come first (nAm, nBm of them):
    // orbits of a single-member alpha and beta class (lone determinants, no
    // operator acts) are the rectangle [nAm, nAc) x [nBm, nBc), left out of
    // the orbit loop and only copied (out of place)
    int nAm, nBm;
    const DevOp *ops;
    int k0, k1, reverse;
    const unsigned long long *prog;
    int split;  // operators k0 .. k0+split-1 act on groups (0, 0), the rest on (1, 1)
    // forward, reachable amplitudes only (list != nullptr): the orbits
    // list[0 .. nlist) and the lone determinants llist[0 .. nllist) (indices
    // into the orbit and lone index spaces); the others hold zero on input
    // and output and are not touched.  Adjoint with stored states (LV
    // kernel): list[0 .. nlist) = the orbits where the stored forward state
    // can be nonzero; the others are skipped (their brackets are zero and
    // their lambda is not used by any later segment)
    const int *list, *llist;
    int nlist, nllist;
};
// BR = adjoint sweep: operators in reverse order with brackets (a.reverse is
// ignored); !BR: forward order.  BC (forward, out of place): v0 is one plain
// vector of N amplitudes read by every lane (the initial state), not a lane-
// layout chunk.
// LV (adjoint): a.list holds the orbits to act on.
template <bool BR, bool BC = false, bool LV = false>
__global__ void k_orb2v(Orb2Args a, const double *v0, double *xout, double *__restrict__ v1, long long cstride,
                        const double2 *__restrict__ cs, int M, double *__restrict__ partial, int slots);
// The same two-group segments on vector-major vectors (grid.y = vector,
// vstride apart; angle table cs[v*M + k]; in place): thread per orbit, its
// up to 16 members gathered into registers.  Each thread's orbit has its own
// type, so the operator words differ across a wave.  BR: v1 holds lambda,
// rotated alike; each operator's bracket is reduced per block into
// partial[(v*M + k)*slots + block] (grid.x <= slots).
template <bool BR>
__global__ void k_orb2s(Orb2Args a, double *v0, double *v1, long long vstride, const double2 *__restrict__ cs, int M,
                        double *__restrict__ partial, int slots);
// Hessian forward-tangent step over a two-group segment, lane layout over
// tangents (chunk c = blockIdx.y holds tangents 64c .. 64c+63 of B, cstride
// apart): psi and lambda at the segment's start (plain vectors, the same for
// every lane; not written) are rotated operator by operator in registers,
// and after operator k's rotation (angle table cs[k]) the tangents rotate,
// the tangent at position inj[k] gains scale[k] G_k psi, and the brackets
// X[k][tangent] = <lambda|G_k|B_tangent> and D[k] = <lambda|G_k G_k|psi> are
// accumulated: xpart[((c*M + k)*slots + block)*64 + lane],
// dpart[k*slots + block] (chunk 0).  grid.x <= slots.
struct TanArgs {
    const double *psi, *lam;
    const int *inj;        // tangent position of operator k's parameter
    const double *scale;   // scale of operator k
    double *xpart, *dpart;
    int slots;
};
__global__ void k_orb2t(Orb2Args a, TanArgs t, double *B, long long cstride, const double2 *__restrict__ cs, int M);
// Three-group orbit segment in lane layout: three consecutive runs of
// operators on pairwise disjoint orbital groups (three TUPS units on
// disjoint spatial pairs).  Classes have up to 8 members per spin; ranks are
// two int4 per class (group-3 member 0, then 1; within each, group-1 member
// + 2 group-2 member) and info = spectator bits | shape1 << 24 | shape2 << 26
// | shape3 << 28.  Classes with two group-3 members come first (nA2, nB2 of
// them).  Groups 1 and 2 are held in registers as in k_orb2v (16 members);
// the group-3 members of an orbit sit in 1, 2 or 4 lanes of one vector (so a
// wave covers 64, 32 or 16 vectors of it), and the third run's operators
// rotate across those lanes.  Work items: 4 per orbit with two group-3
// members in both spins (16 vectors each), 2 with two in one spin, 1
// otherwise; w[0..3] are the cumulative ends of the four item ranges.
// prog[type12 * kOrb3MaxOps + kk] are run 0/1 words as in k_orb2v (type12 from
// shapes 1 and 2); prog3[type3 * kOrb3MaxOps + kk] are the run-2 words over the
// four group-3 slots (a3 + 2 b3), type3 = shapeA3 + 3 shapeB3.  The host
// checks that no operator's sign mask covers another group's orbitals and
// that the active pairs of a run-2 operator share one slot xor.
constexpr int kOrb3MaxOps = 16;
struct Orb3Args {
    int n, nb;
    const int4 *aRanks, *bRanks;
    const uint32_t *aInfo, *bInfo;
    int nA, nB, nA2, nB2;
    long long w[4];
    const DevOp *ops;
    int k0, k1;
    const unsigned long long *prog, *prog3;
    int s1, s2;  // operators k0 .. k0+s1-1: run 0; .. k0+s2-1: run 1; the rest run 2
};
template <bool BR>
__global__ void k_orb3v(Orb3Args a, const double *v0, double *xout, double *__restrict__ v1, long long cstride,
                        const double2 *__restrict__ cs, int M, double *__restrict__ partial, int slots);
// one operator on whole lane-layout vectors (any sector), wave per amplitude
template <bool BR>
__global__ void k_opv(double *__restrict__ v0, double *__restrict__ v1, long long cstride, long long N, DevSector sec,
                      const DevOp *__restrict__ ops, int k, const double2 *__restrict__ cs, int M,
                      double *__restrict__ partial, int slots);
// [v][N] -> lane layout (lanes >= nvec zero-filled) and back; grid (ceil(N/64), chunks)
__global__ void k_to_lanes(const double *__restrict__ src, double *__restrict__ dst, long long N, int nvec);
__global__ void k_from_lanes(const double *__restrict__ src, double *__restrict__ dst, long long N, int nvec);
// dst (lane layout, nchunks chunks) = src in every lane
__global__ void k_bcast_lanes(const double *__restrict__ src, double *__restrict__ dst, long long N, int nchunks);
// out[v*M + k] = scale * sum of operator k's partial rows (rows[k] of them,
// every row when rows is null; row stride slots); grid chunks*M
__global__ void k_reduce_lanes(const double *__restrict__ part, int M, int slots, const int *__restrict__ rows,
                               int nvec, double *__restrict__ out, double scale);

// One operator on whole vectors (any sector); BR = also rotate v1 and
// accumulate the adjoint bracket s*(v1_b v0_b' - v1_b' v0_b) before the
// rotation, one partial per block.  grid.y = independent vector batch.
template <bool BR>
__global__ void k_op_generic(double *__restrict__ v0, double *__restrict__ v1, long long vstride, long long N,
                             DevSector sec, const DevOp *__restrict__ ops, int k, const double2 *__restrict__ cs,
                             long long csStride, double *__restrict__ partial, int slots, long long pstride);

// Complex amplitudes are held as two planes of N doubles, the real plane
// then the imaginary one; the excitation operators rotate both planes
// alike.  A number operator (cre == ann in its DevOp) acts on a basis state
// b only when every orbital it names is occupied and its ordering phase
// (flip) is +1; there it multiplies the amplitude by exp(+i theta)
// (RunFuseNDiagonal, src/fusedevolve.cpp:627-717).  Runs of consecutive
// number operators commute and are applied in one pass of k_phase.
constexpr int kPhaseMaxOps = 16;

__device__ __forceinline__ bool number_op_active(const DevOp &o, uint32_t b)
{
    return o.flip == 1 && (b & o.ann) == o.ann;
}

// Number operators k0 .. k1-1 (k1 - k0 <= kPhaseMaxOps) on complex vectors
// v0 (planes at v0 + 2 c vstride and + vstride for complex vector c =
// blockIdx.y), with the angle table of plane 2c.  BR: v1 holds lambda
// (same layout), rotated alike, and each operator's adjoint bracket
// sum_b Im(lambda_b) Re(v0_b) - Re(lambda_b) Im(v0_b) over its active
// states goes to partial[2c pstride + k slots + block] (a common phase
// leaves the bracket unchanged, so it is taken before the rotation).
template <bool BR>
__global__ void k_phase(double *__restrict__ v0, double *__restrict__ v1, long long vstride, long long N,
                        const uint32_t *__restrict__ states, const DevOp *__restrict__ ops, int k0, int k1,
                        const double2 *__restrict__ cs, long long csStride, double *__restrict__ partial, int slots,
                        long long pstride);

// A fused segment of operators on one tile held in LDS (S_z sector).
template <int NV, bool BR>
__global__ void k_seg(SegArgs a, double *__restrict__ v0, double *__restrict__ v1, long long vstride,
                      const double2 *__restrict__ cs, long long csStride, double *__restrict__ partial, int slots,
                      long long pstride);

// y = M^T x with uniform-width rows stored entry-major (col/val[e*N + j]);
// V vectors per thread share each (col, val) load; per-block partials of
// sum_j x_j y_j for the energy.
template <int V>
__global__ void k_ell(const int32_t *__restrict__ col, const double *__restrict__ val, int W, long long N,
                      const double *__restrict__ x, double *__restrict__ y, long long vstride,
                      double *__restrict__ epart, int eslots);

// y = M^T x with variable rows (rowptr/col/val), same energy partials.
__global__ void k_rows(const int64_t *__restrict__ rowptr, const int32_t *__restrict__ col,
                       const double *__restrict__ val, long long N, const double *__restrict__ x,
                       double *__restrict__ y, long long vstride, double *__restrict__ epart, int eslots);

// out[r] = scale * sum_c part[r*rstride + c*cstride], c < ncols; block per row.
__global__ void k_reduce(const double *__restrict__ part, int nrows, int ncols, long long rstride,
                         long long cstride, double *__restrict__ out, double scale);

// Uniform-width rows of the S_z-sector Hamiltonian from integrals.
struct EllBuildArgs {
    int n, ku, kd, nb;
    long long N;
    int W;
    const uint32_t *unA, *unB;
    const int32_t *rkA, *rkB;
    const double *h, *g;
    int32_t *col;
    double *val;
};
__global__ void k_ell_build(EllBuildArgs a);

// Matrix-free sigma build (S_z sector) from integrals, string driven: one
// workgroup per (alpha string, vector), threads over beta strings.  Tables
// built once per Hamiltonian by k_mf_tables:
//   diag[N]            diagonal elements,
//   FB[t*nb + b]       beta single t of string b: its sign times the one-body
//                      term plus its same-spin two-body spectator terms,
//   FA[t*na + a]       the same for alpha strings.
// Single t of a string enumerates r over occupied orbitals (ascending) and q
// over empty ones (ascending): t = ir * (n - k) + iq.
struct MfArgs {
    int n, ku, kd, na, nb;
    long long N;
    const uint32_t *unA, *unB;
    const int32_t *rkA, *rkB;
    const double *h, *g;
    double *diag, *FA, *FB;
};
__global__ void k_mf_tables(MfArgs a);
__global__ void k_mf_sigma(MfArgs a, const double *__restrict__ x, double *__restrict__ y, long long vstride,
                           double *__restrict__ epart);

// Matrix-free sigma in lane layout (batches, n <= 12, N_down <= 924):
//   y = sparse part (k_sigw_sparse: alpha-beta doubles, the beta-diagonal
//       parts of alpha singles, the alpha-dependent parts of beta singles,
//       the diagonal)
//     + V_bb applied to every alpha row (DGEMM, beta singles' alpha-free
//       parts and beta-beta doubles)
//     + V_aa applied across alpha rows (DGEMM, alpha singles' beta-free parts
//       and alpha-alpha doubles).
constexpr int kSigMaxNb = 924;
constexpr int kSigNN = 144;        // coefficient slots per sign (n <= 12)
constexpr int kSigGS = 2 * kSigNN + 8;  // coefficient row: +coef[e], -coef[e] at kSigNN + e, zeros from 2 kSigNN
struct SigvArgs {
    MfArgs m;
    const uint32_t *ts;  // nb x ntp beta entries (k_beta_table)
    // k_sigg_sparse: per alpha row its source rows in ascending rank,
    // {J, rA, qA, 0} (rA = -1: the identity row), 1 + k(n - k) per row
    const int4 *src;
    int ntp;
    const double *x;
    double *y;
    long long cstride;   // N * 64
    // k_sigg_sparse, exchange-symmetric batches: z = V_aa x (same layout as
    // x); y[a][b] = the sums + z[a][b] + z[b][a], stored for the targets
    // b >= a at least (a wave skips a task whose rows all lie below a);
    // k_mirror_lanes fills b < a.  nullptr: every target, y = the sums
    const double *z;
};
__global__ void k_dense_spin(MfArgs a, int spin, double *__restrict__ V);
// entries: source byte offset J * xunit | coefficient byte offset (slot
// e + parity * nnoff) << cshift
__global__ void k_beta_table(MfArgs a, uint32_t *__restrict__ ts, int ntp, int xunit, int nnoff, int cshift,
                             int sepdiag);
// 16 vectors per workgroup, one staged source row; table with xunit 128,
// cshift 17, kSigWNtp entries per beta and the diagonal separate (SigvArgs
// ntp = singles per beta rounded up to 4)
constexpr int kSigWVC = 16;
constexpr int kSigWThreads = 1024;
constexpr int kSigWNI = (kSigMaxNb + 63) / 64;  // target passes (64 betas per pass)
constexpr int kSigWPre = (kSigMaxNb * kSigWVC + kSigWThreads - 1) / kSigWThreads;
constexpr int kSigWH = 3;                       // entry chunks of 16 (n <= 12: at most 42 entries)
constexpr int kSigWNtp = 16 * kSigWH;
constexpr int kSigWSlots = (kSigMaxNb + 7) / 8 * 8;  // staged betas (copied 8 at a time)
constexpr size_t kSigWLds = (size_t)(kSigGS + 128 + kSigWSlots * kSigWVC) * sizeof(double);
// Two-half variant (n = 2 kd): the source row split by the occupation of
// beta orbital kSigHOrb, halves double-buffered in LDS.  Target slots: the
// strings with that orbital occupied (half A, nA of them, padded to TA, a
// multiple of 4), then the others (half B) up to T.  Table row per target
// slot (kSigWNtp entries): its entries for the A phase, then for the B phase
// (A targets: S same-half then X cross; B targets: X cross then S same;
// entry = source slot * 128 | coefficient byte offset << 17), and in the
// last three slots the beta rank (0xffffffff for padding), the byte offset
// of its own slot in its half, and the string.  hl[h * kSigHMax + slot] =
// beta rank of a half's slot.
constexpr int kSigHOrb = 0;
constexpr int kSigHMax = 464;  // slots per half buffer (C(11, 5) = 462, padded to 8)
struct SighArgs {
    MfArgs m;
    const uint32_t *ts;
    const int *hl;
    int nA, nB, TA, T;
    const double *x;
    double *y;
    long long cstride;  // N * 64
};
constexpr int kSigHWaves = kSigWThreads / 64 - 1;  // summing waves (the last wave stages)
constexpr int kSigHPass = 4 * kSigHWaves;              // targets per pass
constexpr int kSigHNI = (2 * kSigHMax + kSigHPass - 1) / kSigHPass;
constexpr size_t kSigHLds =
    (size_t)(2 * kSigHMax * kSigWVC + 2 * kSigGS + 2 * 128) * sizeof(double) + 2 * kSigHMax * sizeof(int);
template <int S, int X>
__global__ void k_sigh_sparse(SighArgs a);  // grid (ceil(na / 8) * 32, chunks), kSigWThreads, kSigHLds
// NG = groups of 4 single entries per beta (1..9; a.ntp = 4 NG)
template <int NG>
__global__ void k_sigw_sparse(SigvArgs a);  // grid (ceil(na / 8) * 32, chunks), kSigWThreads, kSigWLds
// 12 orbitals, K beta electrons (K = 6): targets grouped by their upper
// six orbitals, one group per 16-lane row (sigg.hip); same grid, block and LDS
template <int K>
__global__ void k_sigg_sparse(SigvArgs a);
// y[a][b] = y[b][a] for b < a (lane layout, nb strings per spin, cstride
// doubles per chunk); grid (nb, chunks), kBlock
__global__ void k_mirror_lanes(double *__restrict__ y, int nb, long long cstride);
__global__ void k_dot_lanes(const double *__restrict__ x, const double *__restrict__ y, long long N,
                            double *__restrict__ part);  // grid (G, chunks)

// The same split for vector-major vectors (n <= 14): sparse part per (alpha
// row, vector), same-spin parts as DGEMMs over whole vectors.
constexpr int kSig1Threads = 1024;
constexpr int kSig1MaxNb = 3432;  // C(14, 7)
constexpr int kSig1BI = (kSig1MaxNb + kSig1Threads - 1) / kSig1Threads;
constexpr int kSig1NN = 196;      // n <= 14
constexpr int kSig1GS = 2 * kSig1NN + 8;
constexpr size_t kSig1Lds = (size_t)(2 * kSig1GS + 2 * kSig1MaxNb) * sizeof(double);
struct Sig1Args {
    MfArgs m;
    const uint32_t *ts;  // nb x ntp entries, xunit 8, nnoff kSig1NN
    int ntp;
    const double *x;
    double *y;
    long long vstride;
    // k_sig1r: 1 for an exchange-symmetric vector, whose targets b < a are
    // left to k_mirror_rows; 0 every target
    int half;
};
__global__ void k_sig1_sparse(Sig1Args a);
// y[a][b] = y[b][a] for b < a, vector-major (nb strings per spin, vstride
// doubles per vector), through 32 x 32 LDS tiles; grid (ceil(nb / 32),
// ceil(nb / 32), nvec), kBlock
__global__ void k_mirror_rows(double *__restrict__ y, int nb, long long vstride);  // grid (na, nvec), kSig1Threads, kSig1Lds
// RB source rows per phase; LDS RB * (kSig1GS + nb) doubles
template <int RB>
__global__ void k_sig1r(Sig1Args a);
__global__ void k_dot_rows(const double *__restrict__ x, const double *__restrict__ y, long long N, long long vstride,
                           double *__restrict__ part);  // grid (G, nvec)


// ------------------------------------------------- window passes (lane layout)
// A window is a set of w spatial orbitals (the same bits in the alpha and in
// the beta string).  Determinants with the same bits outside the window in
// both strings form a tile: (alpha outer class, beta outer class), members =
// the inner patterns of the class's inner electron counts, C(w, ca) x
// C(w, cb).  Every orbit segment (phase) whose bits lie inside the window
// maps each tile into itself, so one pass stages a tile of V vectors in LDS
// and applies all of the pass's phases there: a sweep reads and writes the
// state once per pass instead of once per segment.
constexpr int kWinMaxW = 8;        // window width
constexpr int kWinMaxOps = 64;     // operators per pass
constexpr int kWinMaxPhases = 32;  // phases per pass
struct WinPhaseDev {
    int k0, nk;       // operators k0 .. k0 + nk - 1 (program order)
    int progOff;      // orbit program words at prog[progOff + type * kOrbMaxOps + kk] (within the pass)
    int kp;           // position of operator k0 in the pass's operator list
    uint32_t ua, ub;  // the phase's alpha / beta bits (string level)
    uint32_t spect;   // full-string mask of its spectator bits
    uint32_t pad;
};
// Tables of one pass, one blob of 32-bit words in device memory, read by
// the kernel through the scalar unit (every value is wave-uniform); word
// offsets:
//   0       phases (WinPhaseDev, 8 words each)
//   oSubIdx sub-class index, 4 words per (phase, spin, inner count) at
//           (phase * 2 + spin) * (w + 1) + count: {offset in sub, n | n2 << 16,
//           magic(n), magic(n2)}, n sub-classes of which the first n2 have two
//           members, magic(d) = ceil(2^24 / d) (x / d = (x * magic) >> 24 for
//           x < 2^24 / d)
//   oSub    sub-class entries: i0 | i1 << 8 (0xff: none) | shape << 16, the
//           inner patterns (indices within their count) agreeing outside
//           the phase's bits; shape as in OrbArgs
//   oProg   orbit programs of the phases (9 types x kOrbMaxOps words each)
//   oMask   sign masks of the pass's operators, in pass order
//   oPat    placed inner patterns by (count, value), 2^w
//   oCOff   w + 2 offsets of the counts in oPat
//   oCls    outer classes, alpha then beta: outer bits | inner count << 24
//   oCStart first member of each class in rows (alpha) / rows + nrowsA (beta)
//   oProgT  the programs again by (phase, operator): 16 words per operator,
//           its 9 type words first (16-word aligned), then its phase as a
//           sign bit (0x80000000) when every pair it rotates has the same one
struct WinArgs {
    int n, w, w0;
    const uint32_t *blob;
    int oSubIdx, oSub, oProg, oMask, oPat, oCOff, oCls, oCStart, oProgT;
    // members of the outer classes by class: alpha rank * nb (nrowsA words),
    // then beta ranks
    const uint32_t *rows;
    int nrowsA;
    const int *tiles;  // tiles to process: alpha class * nBcls + beta class
    int ntiles, nAcls, nBcls;
    int nph, nkp;
    const int *kpOp;   // operator of each pass position
    int maxMembers;    // largest tile of the pass
};
// Dynamic LDS of k_win for a pass (bytes): the tile of 64 vectors (psi, and
// lambda in the adjoint), the adjoint's per-wave bracket rows of one phase
// and the pass's bracket rows.
inline size_t win_lds_bytes(bool br, int waves, int maxMembers, int nkp)
{
    size_t b = (size_t)maxMembers * 64 * sizeof(double) * (br ? 2 : 1);
    if (br) b += ((size_t)waves * kOrbMaxOps + (size_t)nkp) * 64 * sizeof(double);
    return b;
}
// One pass over the listed tiles, grid (blocks, chunks), NW waves per block,
// lane = vector of the 64-vector chunk; block b takes the tiles b + j * grid.
// Forward (BR false): in -> out (may be equal); BC: in is one plain vector
// broadcast to every lane.  Adjoint (BR): phases and operators in reverse
// with the negated angles, psi from in (written back to out unless null),
// lambda in place, per-operator brackets into partial rows (slot = b).
template <bool BR, bool BC, int NW>
__global__ void k_win(WinArgs a, const double *in, double *out, double *lam, long long cstride,
                      const double2 *__restrict__ cs, int M, double *__restrict__ partial, int slots);

// Window passes over four consecutive orbitals whose phases are TUPS units
// (the same adjacent pair in both spins): one wave holds a whole tile (up to
// 6 x 6 members for 64 vectors) in registers, the member indices of every
// sub-orbit fixed at compile time per pair of inner counts, so no LDS or
// barrier is needed between phases.  Grid-stride over the listed tiles by
// wave (NW waves per block); the forward stages the pass's angles in LDS
// (nkp x 64 double2); the adjoint keeps its bracket rows per wave in LDS
// (nkp x 64 doubles each) and writes the block's sum as its partial row
// (slot = block).  Arguments as k_win, with WinArgs::w0 the window's first
// orbital.
// TU: every phase is a TUPS unit of five operators (beta single, alpha
// single, paired double, beta single, alpha single) whose program words
// rotate exactly the member pairs of that pattern (checked on the host).
// epart (adjoint, or null): each block also writes the sum of x * l over its
// waves' tiles before their first phase, per lane, as row b / P of row group
// b % P of chunk c (P = 2^min(4, trailing zeros of gridDim.x), gridDim.x / P
// rows per group): launched on the sweep's first adjoint pass, where x is
// the final state and l = H x, the rows sum to the energies.
template <bool BR, bool BC, int NW, bool TU>
__global__ void k_winr(WinArgs a, const double *in, double *out, double *lam, long long cstride,
                       const double2 *__restrict__ cs, int M, double *__restrict__ partial, int slots,
                       double *__restrict__ epart);

// Window passes in the vector-major layout (one vector per grid row,
// vstride doubles apart), register tiles of four-orbital windows as
// k_winr, lane = tile: a wave takes a task {class = 5 ca + cb, offset of
// its record block in rec, count <= 64} of tiles of one class of inner
// counts; the block holds (NA + NB + 2) words per lane, word q at
// [q * 64 + lane]: the tile's NA alpha rows (alpha rank * nb), its NB beta
// rows (beta rank), its alpha and beta outer bits.  Forward and adjoint in
// place (psi is written back in the adjoint); the adjoint sums its brackets
// per block (4 waves, LDS rows of nkp x 64 per wave) into
// partial[(vector * M + k) * slots + block].
template <bool BR, bool TU>
__global__ void k_winv(WinArgs a, const int4 *__restrict__ tasks, int ntasks, const uint32_t *__restrict__ rec,
                       double *v, double *lam, long long vstride, const double2 *__restrict__ cs, int M,
                       double *__restrict__ partial, int slots);

// Window passes in the vector-major layout with staged alpha rows: a block
// takes one alpha outer class at a time (grid-stride), copies its alpha rows
// of the vector (at most kWinVsRows of nb doubles; the adjoint also
// lambda's) to LDS with coalesced loads, and each of its waves takes tasks
// {beta inner count, record block offset, count <= 64} of beta outer
// classes (lane = tile), whose record block holds per lane word q at
// [q * 64 + lane]: NB beta rows (beta ranks), then the beta outer bits; the
// tile's members are read from and written back to the staged rows, which
// go back to the vector after the class.  Brackets as k_winv (per block:
// partial[(vector * M + k) * slots + block]).  LDS: kWinVsRows x nb doubles
// (twice for the adjoint) + the adjoint's waves x nkp x 64 doubles.
constexpr int kWinVsRows = 6;
template <bool BR, bool TU>
__global__ void k_winvs(WinArgs a, const int4 *__restrict__ tasks, int ntasks, const uint32_t *__restrict__ rec,
                        int nb, double *v, double *lam, long long vstride, const double2 *__restrict__ cs, int M,
                        double *__restrict__ partial, int slots);

// Hessian tangent pass over four-orbital windows whose phases are TUPS
// units: lane = tangent (64 per chunk, grid.y = chunks, in place in B, lane
// layout), the tile's psi and lambda (one vector each, psiIn/lamIn) in
// registers beside it.  Every operator rotates all three, its injection lane
// (inj[k] = chunk * 64 + lane) adds scale[k] G psi, and the brackets
// X[lane][k] = <lambda|G_k|B_lane> and D[k] = <lambda|G_k G_k|psi> are summed
// per block into xpart[((c * M + k) * slots + block) * 64 + lane] and (chunk
// 0) dpart[k * slots + block].  Chunk 0 writes psi and lambda after the pass
// to psiOut/lamOut.  a.tiles lists every tile of the pass, the nreach that
// psi can have reached first; on the others psi and the tangents are zero
// and stay zero, so chunk 0 carries lambda there and writes psi's zeros.
// 2 waves per block, LDS 2 x nkp x 65 doubles.
__global__ void k_wint(WinArgs a, const double *psiIn, const double *lamIn, double *psiOut, double *lamOut, double *B,
                       long long cstride, const double2 *__restrict__ cs, int M, const int *__restrict__ inj,
                       const double *__restrict__ scale, double *__restrict__ xpart, double *__restrict__ dpart,
                       int slots, int nreach);

// Hessian tangent step for operator k (see Engine::hessian).
template <int MU>
__global__ void k_tangent(const double *__restrict__ psiIn, double *__restrict__ psiOut,
                          const double *__restrict__ lamIn, double *__restrict__ lamOut, double *__restrict__ B,
                          long long bstride, int nactive, long long N, DevSector sec,
                          const DevOp *__restrict__ ops, int k, const double2 *__restrict__ cs, int injectMu,
                          double scalek, double *__restrict__ xpart, int G, double *__restrict__ dpart);

// part[chunk][i][j] = sum over the chunk of A[i][.] * B[j][.]
__global__ void k_gram(const double *__restrict__ A, const double *__restrict__ B, long long stride, int P,
                       long long N, int chunks, double *__restrict__ part);

} // namespace ae
