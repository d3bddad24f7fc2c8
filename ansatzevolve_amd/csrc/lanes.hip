// Lane-layout sweep kernels (batches of 64 vectors, lane = vector): orbit
// segments (k_orbv, k_orb2v, k_orb3v) and one-operator passes (k_opv).  See
// kernels.hpp.
//
// Built without the promotion of private arrays to vector registers
// (-mllvm -disable-promote-alloca-to-vector): the 16-member orbit arrays
// would otherwise become one 32-register tuple whose every conditional
// update copies the whole tuple (thousands of register moves and 150+
// VGPRs in the adjoint two-group kernel instead of 95).
#include "kernels.hpp"

namespace ae {

// Read-only tables indexed by wave-uniform values are read through the
// constant address space, so they come through the scalar unit.  Through
// generic pointers the compiler cannot rule out that the kernel's own stores
// (no restrict: the input and output state may be one buffer) change them,
// loads them as vector loads and waits for every outstanding store first.
__device__ __forceinline__ uint32_t ld_const(const uint32_t *p, int i)
{
    return ((const __attribute__((address_space(4))) uint32_t *)p)[i];
}
__device__ __forceinline__ int4 ld_const(const int4 *p, int i)
{
    const __attribute__((address_space(4))) int *q = (const __attribute__((address_space(4))) int *)p + 4 * i;
    return make_int4(q[0], q[1], q[2], q[3]);
}

// ------------------------------------------------ lane layout (batches)
// Per-wave bracket sums of a segment's operators -> one partial row of 64
// lanes per (operator, block).
template <int NK>
__device__ __forceinline__ void lanes_block_partials(const double (&acc)[NK], int nk, int k0, int k1, int reverse,
                                                     double *red, double *__restrict__ partial, int c, int M,
                                                     int slots)
{
    constexpr int WPB = kBlock / kLanes;
    const int lane = threadIdx.x & (kLanes - 1), wv = threadIdx.x / kLanes;
#pragma unroll
    for (int kk = 0; kk < NK; kk++)
        if (kk < nk) red[(kk * WPB + wv) * kLanes + lane] = acc[kk];
    __syncthreads();
    for (int e = threadIdx.x; e < nk * kLanes; e += kBlock) {
        const int kk = e / kLanes, l = e - kk * kLanes;
        double s = 0;
#pragma unroll
        for (int w = 0; w < WPB; w++) s += red[(kk * WPB + w) * kLanes + l];
        const int k = reverse ? k1 - 1 - kk : k0 + kk;
        partial[(((long long)c * M + k) * slots + blockIdx.x) * kLanes + l] = s;
    }
}

// One orbit of a lane-layout wave: member indices, presence, type and the
// spectator bits are wave-uniform; amplitudes are per lane (per vector).
struct LaneOrbit {
    long long idx[4];
    bool ok[4];
    bool live;  // members are loaded and stored
    bool act;   // operators act (more than one member)
    int type;
    uint32_t spect;
};

__device__ __forceinline__ void lane_orbit_open(const OrbArgs &a, int t, LaneOrbit &o, bool copyAll)
{
    const int ia = t / a.nBc, ib = t - ia * a.nBc;
    const int4 da = ld_const(a.aDesc, ia), db = ld_const(a.bDesc, ib);
    const int Ra = da.y >= 0 ? 2 : 1, Rb = db.y >= 0 ? 2 : 1;
    o.act = Ra * Rb > 1;  // a lone determinant: no operator of the segment acts on it
    o.live = o.act || copyAll;
    const int ta = Ra == 2 ? 0 : (((uint32_t)da.z & a.uaLocal) ? 2 : 1);
    const int tb = Rb == 2 ? 0 : (((uint32_t)db.z & a.ubLocal) ? 2 : 1);
    o.type = ta * 3 + tb;
    o.spect = (((uint32_t)da.z << a.n) | (uint32_t)db.z) & a.spect;
    const int ra[2] = {da.x, da.y >= 0 ? da.y : da.x};
    const int rb[2] = {db.x, db.y >= 0 ? db.y : db.x};
#pragma unroll
    for (int p = 0; p < 4; p++) {
        const int i = p >> 1, j = p & 1;
        o.ok[p] = o.live && (i < Ra) && (j < Rb);
        o.idx[p] = ((long long)ra[i] * a.nb + rb[j]) * kLanes;
    }
}

template <bool BR>
__device__ __forceinline__ void lane_orbit_apply(const OrbArgs &a, int nk, const LaneOrbit &o, double (&x)[4],
                                                 double (&l)[4], const double (&C)[kOrbMaxOps],
                                                 const double (&S)[kOrbMaxOps], double (&acc)[kOrbMaxOps])
{
    constexpr int PP[6] = {0, 0, 0, 1, 1, 2}, QQ[6] = {1, 2, 3, 2, 3, 3};
    const uint32_t *prog = a.prog + o.type * kOrbMaxOps;
#pragma unroll
    for (int kk = 0; kk < kOrbMaxOps; kk++) {
        if (kk >= nk) break;
        const int k = a.reverse ? a.k1 - 1 - kk : a.k0 + kk;
        const uint32_t w = prog[a.reverse ? nk - 1 - kk : kk];
        if (!(w & 0x3f)) continue;
        const bool neg = __popc(o.spect & a.ops[k].mask) & 1;
        double br = 0;
#pragma unroll
        for (int r = 0; r < 6; r++) {
            if (!((w >> r) & 1)) continue;
            const int p = PP[r], q = QQ[r];
            const double s = (((w >> (8 + r)) & 1) != neg) ? -1.0 : 1.0;
            const double xp = x[p], xq = x[q];
            if (BR) {
                const double lp = l[p], lq = l[q];
                br += s * (lp * xq - lq * xp);
                l[p] = s * lq * S[kk] + lp * C[kk];
                l[q] = -s * lp * S[kk] + lq * C[kk];
            }
            x[p] = s * xq * S[kk] + xp * C[kk];
            x[q] = -s * xp * S[kk] + xq * C[kk];
        }
        acc[kk] += br;
    }
}

// Orbit segment in lane layout.  Each wave takes two orbits per iteration
// and issues the loads of both before computing either, so that twice the
// bytes are in flight per wave (the kernel is bound by HBM latency x bytes
// in flight, not by arithmetic).
template <bool BR>
__global__ void __launch_bounds__(kBlock)
    k_orbv(OrbArgs a, const double *v0, double *xout, double *__restrict__ v1, long long cstride,
           const double2 *__restrict__ cs, int M, double *__restrict__ partial, int slots)
{
    constexpr int WPB = kBlock / kLanes;
    __shared__ double red[BR ? kOrbMaxOps * WPB * kLanes : 1];
    const int lane = threadIdx.x & (kLanes - 1);
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / kLanes);
    const int c = blockIdx.y;
    v0 += c * cstride + lane;
    if (xout) xout += c * cstride + lane;
    if (BR) v1 += c * cstride + lane;
    const bool copyAll = xout && xout != v0;  // out of place: lone determinants are copied too
    cs += (long long)c * M * kLanes + lane;
    const int nk = a.k1 - a.k0;
    double C[kOrbMaxOps], S[kOrbMaxOps], acc[kOrbMaxOps];
#pragma unroll
    for (int kk = 0; kk < kOrbMaxOps; kk++) {
        acc[kk] = 0;
        C[kk] = 1;
        S[kk] = 0;
        if (kk < nk) {
            const int k = a.reverse ? a.k1 - 1 - kk : a.k0 + kk;
            const double2 t = cs[(long long)k * kLanes];
            C[kk] = t.x;
            S[kk] = t.y;
        }
    }
    const int total = a.nAc * a.nBc;
    const int stride = gridDim.x * WPB;
    for (int t = blockIdx.x * WPB + wv; t < total; t += 2 * stride) {
        LaneOrbit o0, o1;
        lane_orbit_open(a, t, o0, copyAll);
        const bool has1 = t + stride < total;
        if (has1) lane_orbit_open(a, t + stride, o1, copyAll);
        else o1.live = o1.act = false;
        if (!o0.live && !o1.live) continue;
        if (!o1.live) {
#pragma unroll
            for (int p = 0; p < 4; p++) o1.ok[p] = false;
        }
        double x0[4], l0[4], x1[4], l1[4];
#pragma unroll
        for (int p = 0; p < 4; p++) {
            x0[p] = o0.ok[p] ? v0[o0.idx[p]] : 0.0;
            x1[p] = o1.ok[p] ? v0[o1.idx[p]] : 0.0;
            if (BR) {
                l0[p] = o0.ok[p] ? v1[o0.idx[p]] : 0.0;
                l1[p] = o1.ok[p] ? v1[o1.idx[p]] : 0.0;
            }
        }
        if (o0.act) lane_orbit_apply<BR>(a, nk, o0, x0, l0, C, S, acc);
        if (o1.act) lane_orbit_apply<BR>(a, nk, o1, x1, l1, C, S, acc);
#pragma unroll
        for (int p = 0; p < 4; p++) {
            if (o0.ok[p]) {
                if (xout) xout[o0.idx[p]] = x0[p];
                if (BR) v1[o0.idx[p]] = l0[p];
            }
            if (o1.ok[p]) {
                if (xout) xout[o1.idx[p]] = x1[p];
                if (BR) v1[o1.idx[p]] = l1[p];
            }
        }
    }
    if (BR) lanes_block_partials<kOrbMaxOps>(acc, nk, a.k0, a.k1, a.reverse, red, partial, c, M, slots);
}

// Two-group orbit segment in lane layout (see kernels.hpp): wave per orbit
// of up to 16 members, lane per vector, the two runs of operators unrolled.
// Uniform operator word: the same 6 pair bits and phases in all four
// sub-orbits; each active pair is rotated in every sub-orbit.
// (p, q) <- (C p + sS q, C q - sS p) with both results written back into
// the registers of p and q.  Written out so that a conditional rotation
// leaves the 16-member arrays where they were: with compiler-chosen
// registers every taken pair block ended in ~32 register copies.
__device__ __forceinline__ void rot_inplace(double &p, double &q, double C, double sS)
{
    double t, u;
    // operands: %0 p, %1 q, %2 %3 temporaries, %4 C, %5 sS
    asm("v_mul_f64 %2, %5, %1\n\t"
        "v_mul_f64 %3, %5, %0\n\t"
        "v_fma_f64 %0, %4, %0, %2\n\t"
        "v_fma_f64 %1, %4, %1, -%3"
        : "+v"(p), "+v"(q), "=&v"(t), "=&v"(u)
        : "v"(C), "v"(sS));
}

// Adjoint step on a pair: brr += lp xq - lq xp (before the step), then the
// rotation of (xp, xq) and of (lp, lq), every result in place.
__device__ __forceinline__ void rot2_br_inplace(double &xp, double &xq, double &lp, double &lq, double &brr,
                                                double C, double sS)
{
    double t2, u2, t3, u3;
    // operands: %0 xp, %1 xq, %2 lp, %3 lq, %4 brr, %5-%8 temporaries,
    // %9 C, %10 sS.  The bracket goes straight into brr by two FMAs on the
    // old values (10 FP64 instructions per pair instead of 11: the sweeps
    // are about as much FP64-issue as memory bound)
    asm("v_fma_f64 %4, %2, %1, %4\n\t"
        "v_mul_f64 %5, %10, %3\n\t"
        "v_mul_f64 %6, %10, %2\n\t"
        "v_mul_f64 %7, %10, %1\n\t"
        "v_mul_f64 %8, %10, %0\n\t"
        "v_fma_f64 %4, -%3, %0, %4\n\t"
        "v_fma_f64 %2, %9, %2, %5\n\t"
        "v_fma_f64 %3, %9, %3, -%6\n\t"
        "v_fma_f64 %0, %9, %0, %7\n\t"
        "v_fma_f64 %1, %9, %1, -%8"
        : "+v"(xp), "+v"(xq), "+v"(lp), "+v"(lq), "+v"(brr), "=&v"(t2), "=&v"(u2), "=&v"(t3), "=&v"(u3)
        : "v"(C), "v"(sS));
}

// cm: the sub-orbits (c = other alpha member + 2 other beta member) whose
// members exist in this orbit type; the others hold only absent (zero)
// amplitudes and are skipped (most orbits have 2 or 4 of 16 members)
template <int G, int H>
__device__ __forceinline__ void orb2_apply_uniform(unsigned w, bool neg, double C, double S, double (&x)[16],
                                                   double (&l)[16], bool br_on, double &br, unsigned cm)
{
    constexpr int PP[6] = {0, 0, 0, 1, 1, 2}, QQ[6] = {1, 2, 3, 2, 3, 3};
#pragma unroll
    for (int r = 0; r < 6; r++) {
        if (!((w >> r) & 1u)) continue;
        const double s = ((((w >> (6 + r)) & 1u) != 0) != neg) ? -1.0 : 1.0;
        const double sS = s * S;
        double brr = 0;
        const int ip = PP[r] >> 1, jp = PP[r] & 1, iq = QQ[r] >> 1, jq = QQ[r] & 1;
#pragma unroll
        for (int c = 0; c < 4; c++) {
            if (c > 0 && !((cm >> c) & 1u)) continue;
            const int oa = c & 1, ob = c >> 1;
            const int sp = ((ip << G) | (oa << (1 - G))) + 4 * ((jp << H) | (ob << (1 - H)));
            const int sq = ((iq << G) | (oa << (1 - G))) + 4 * ((jq << H) | (ob << (1 - H)));
            if (br_on) rot2_br_inplace(x[sp], x[sq], l[sp], l[sq], brr, C, sS);
            else rot_inplace(x[sp], x[sq], C, sS);
        }
        if (br_on) br += s * brr;
    }
}

// one operator of a run: its word, angle (per lane, from LDS) and bracket slot
template <bool BR, int G, int H>
__device__ __forceinline__ void orb2_op(const Orb2Args &a, const unsigned long long *pw, int kk, uint32_t spect,
                                        const double2 (*csL)[kLanes], double *accW, int lane, double (&x)[16],
                                        double (&l)[16], unsigned cmask)
{
    const unsigned w = (unsigned)pw[kk] & 0xfffu;
    if (!w) return;
    const bool neg = __popc(spect & a.ops[a.k0 + kk].mask) & 1;
    const double2 t = csL[kk][lane];
    double br = 0;
    // both sweeps skip the sub-orbits whose members the orbit type lacks
    orb2_apply_uniform<G, H>(w, neg, t.x, t.y, x, l, BR, br, (cmask >> (4 * G)) & 0xfu);
    if (BR) accW[kk * kLanes + lane] += br;
}

template <bool BR, bool BC, bool LV>
// forward: at most 64 VGPRs (8 waves per SIMD, one orbit per wave, absent
// sub-orbits skipped: 6.6 ms per step, against 7.2 ms with two orbits per
// wave, every sub-orbit rotated and 6 waves); adjoint: 96 (5 waves)
__global__ void __launch_bounds__(kBlock, BR ? 5 : 8)
    k_orb2v(Orb2Args a, const double *v0, double *xout, double *__restrict__ v1, long long cstride,
            const double2 *__restrict__ cs, int M, double *__restrict__ partial, int slots)
{
    constexpr int WPB = kBlock / kLanes;
    constexpr int K = kOrb2MaxOps;
    __shared__ double2 csL[K][kLanes];
    __shared__ double accL[BR ? WPB * K * kLanes : 1];
    const int lane = threadIdx.x & (kLanes - 1);
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / kLanes);
    const int c = blockIdx.y;
    const int nk = a.k1 - a.k0, m = a.split;
    for (int e = threadIdx.x; e < nk * kLanes; e += kBlock) {
        const int kk = e / kLanes, l = e - kk * kLanes;
        csL[kk][l] = cs[((long long)c * M + a.k0 + kk) * kLanes + l];
    }
    if (BR)
        for (int e = threadIdx.x; e < WPB * K * kLanes; e += kBlock) accL[e] = 0.0;
    __syncthreads();
    double *accW = BR ? accL + wv * K * kLanes : nullptr;
    if (!BC) v0 += c * cstride;  // BC: v0 is one plain vector, the same for every lane
    if (xout) xout += c * cstride;
    if (BR) v1 += c * cstride;
    // out of place: every member is written, lone determinants included
    const bool copyAll = BC || (xout && xout != v0);
    // orbits with some two-member class: alpha classes [0, nAm) with every
    // beta class, then the single-member alpha classes with beta [0, nBm)
    const int mainA = a.nAm * a.nBc;
    // orbit lists: the forward sweep's reachable orbits; in the LV adjoint
    // the orbits where the stored state can be nonzero.  The adjoint leaves
    // the others alone: their brackets are zero, and the determinants whose
    // lambda later segments combine with a nonzero state (the reachable set
    // before them, a subset of this segment's) lie in the listed orbits, so
    // their lambda is never used again.
    const int *list = (BR && !LV) ? nullptr : a.list;
    const int total = list ? a.nlist : mainA + (a.nAc - a.nAm) * a.nBm;
    const int stride = gridDim.x * WPB;
    // one orbit per wave iteration (two per iteration, loaded together,
    // needed more registers than 8 waves per SIMD allow)
    constexpr int NO = 1;
    for (int t0 = blockIdx.x * WPB + wv; t0 < total; t0 += NO * stride) {
        int rA[NO][4], rB[NO][4], type[NO];
        uint32_t spect[NO], cmask[NO];
        bool live[NO], act[NO];
#pragma unroll
        for (int o = 0; o < NO; o++) {
            const int t = t0 + o * stride;
            live[o] = t < total;
            const int tt = list ? list[live[o] ? t : t0] : (live[o] ? t : t0);
            int ia, ib;
            if (tt < mainA) {
                ia = tt / a.nBc;
                ib = tt - ia * a.nBc;
            } else {
                const int u = tt - mainA, i = u / a.nBm;
                ia = a.nAm + i;
                ib = u - i * a.nBm;
            }
            // adjoint: descriptors through the scalar unit (ld_const); the
            // forward kernel ran slower that way (9.5 against 7.4 ms per step)
            const int4 ra = BR ? ld_const(a.aRanks, ia) : a.aRanks[ia];
            const int4 rb = BR ? ld_const(a.bRanks, ib) : a.bRanks[ib];
            const uint32_t fa = BR ? ld_const(a.aInfo, ia) : a.aInfo[ia];
            const uint32_t fb = BR ? ld_const(a.bInfo, ib) : a.bInfo[ib];
            rA[o][0] = ra.x; rA[o][1] = ra.y; rA[o][2] = ra.z; rA[o][3] = ra.w;
            rB[o][0] = rb.x; rB[o][1] = rb.y; rB[o][2] = rb.z; rB[o][3] = rb.w;
            const int na_ = (ra.x >= 0) + (ra.y >= 0) + (ra.z >= 0) + (ra.w >= 0);
            const int nb_ = (rb.x >= 0) + (rb.y >= 0) + (rb.z >= 0) + (rb.w >= 0);
            act[o] = live[o] && na_ * nb_ >= 2;  // a lone determinant: nothing acts on it
            if (!copyAll) live[o] = act[o];
            type[o] = ((fa >> 24) & 3) + 3 * ((fa >> 26) & 3) + 9 * (((fb >> 24) & 3) + 3 * ((fb >> 26) & 3));
            spect[o] = ((fa & 0xffffffu) << a.n) | (fb & 0xffffffu);
            // existing sub-orbits per run: run 0 (first groups) varies the second
            // groups' members, run 1 the first groups' (shape 0 = two members)
            {
                const bool a2 = ((fa >> 26) & 3) == 0, b2 = ((fb >> 26) & 3) == 0;
                const bool a1 = ((fa >> 24) & 3) == 0, b1 = ((fb >> 24) & 3) == 0;
                cmask[o] = 1u | (a2 ? 2u : 0u) | (b2 ? 4u : 0u) | (a2 && b2 ? 8u : 0u) |
                           ((1u | (a1 ? 2u : 0u) | (b1 ? 4u : 0u) | (a1 && b1 ? 8u : 0u)) << 4);
            }
        }
        // element offsets fit in 32 bits (N * 64 < 2^32 for n <= 14)
        double x[NO][16], l[NO][16];
#pragma unroll
        for (int o = 0; o < NO; o++)
#pragma unroll
            for (int s = 0; s < 16; s++) {
                const int ma = s & 3, mb = s >> 2;
                const bool ok = live[o] && rA[o][ma] >= 0 && rB[o][mb] >= 0;
                const uint32_t off = ((uint32_t)rA[o][ma] * (uint32_t)a.nb + (uint32_t)rB[o][mb]) * kLanes;
                if (BC) x[o][s] = ok ? v0[off / kLanes] : 0.0;
                else x[o][s] = ok ? __builtin_nontemporal_load(v0 + off + lane) : 0.0;
                l[o][s] = (BR && ok) ? __builtin_nontemporal_load(v1 + off + lane) : 0.0;
            }
#pragma unroll
        for (int o = 0; o < NO; o++) {
            if (!act[o]) continue;
            const unsigned long long *pw = a.prog + type[o] * K;
            // the forward sweep (no brackets) runs forward, the adjoint sweep in
            // reverse: a compile-time direction keeps the 16-member arrays in
            // one register assignment (a runtime branch made the compiler copy
            // them at the merge)
            if (!BR) {
#pragma unroll
                for (int kk = 0; kk < K; kk++) {
                    if (kk >= m) break;
                    orb2_op<BR, 0, 0>(a, pw, kk, spect[o], csL, accW, lane, x[o], l[o], cmask[o]);
                }
#pragma unroll
                for (int kk = 0; kk < K; kk++) {
                    if (kk < m || kk >= nk) continue;
                    orb2_op<BR, 1, 1>(a, pw, kk, spect[o], csL, accW, lane, x[o], l[o], cmask[o]);
                }
            } else {
#pragma unroll
                for (int j = 0; j < K; j++) {
                    const int kk = K - 1 - j;
                    if (kk < m || kk >= nk) continue;
                    orb2_op<BR, 1, 1>(a, pw, kk, spect[o], csL, accW, lane, x[o], l[o], cmask[o]);
                }
#pragma unroll
                for (int j = 0; j < K; j++) {
                    const int kk = K - 1 - j;
                    if (kk >= m) continue;
                    orb2_op<BR, 0, 0>(a, pw, kk, spect[o], csL, accW, lane, x[o], l[o], cmask[o]);
                }
            }
        }
#pragma unroll
        for (int o = 0; o < NO; o++)
#pragma unroll
            for (int s = 0; s < 16; s++) {
                const int ma = s & 3, mb = s >> 2;
                if (live[o] && rA[o][ma] >= 0 && rB[o][mb] >= 0) {
                    const uint32_t off = ((uint32_t)rA[o][ma] * (uint32_t)a.nb + (uint32_t)rB[o][mb]) * kLanes;
                    if (xout) __builtin_nontemporal_store(x[o][s], xout + off + lane);
                    if (BR) __builtin_nontemporal_store(l[o][s], v1 + off + lane);
                }
            }
    }
    if (copyAll) {
        // lone determinants: copied, four per wave iteration (all loads first)
        const int nAl = a.nAc - a.nAm, nBl = a.nBc - a.nBm, nl = list ? a.nllist : nAl * nBl;
        for (int t0 = blockIdx.x * WPB + wv; t0 < nl; t0 += 4 * stride) {
            uint32_t off[4];
            double v[4];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int t = t0 + q * stride, tt = list ? a.llist[t < nl ? t : t0] : (t < nl ? t : t0);
                const int i = tt / nBl, ia = a.nAm + i, ib = a.nBm + (tt - i * nBl);
                const int ra = a.aRanks[ia].x, rb = a.bRanks[ib].x;
                off[q] = ((uint32_t)ra * (uint32_t)a.nb + (uint32_t)rb) * kLanes;
                if (t < nl) v[q] = BC ? v0[off[q] / kLanes] : __builtin_nontemporal_load(v0 + off[q] + lane);
            }
#pragma unroll
            for (int q = 0; q < 4; q++)
                if (t0 + q * stride < nl) __builtin_nontemporal_store(v[q], xout + off[q] + lane);
        }
    }
    if (BR) {
        __syncthreads();
        for (int e = threadIdx.x; e < nk * kLanes; e += kBlock) {
            const int kk = e / kLanes, l = e - kk * kLanes;
            double s = 0;
#pragma unroll
            for (int q = 0; q < WPB; q++) s += accL[(q * K + kk) * kLanes + l];
            partial[(((long long)c * M + a.k0 + kk) * slots + blockIdx.x) * kLanes + l] = s;
        }
    }
}

template __global__ void k_orb2v<false, false, false>(Orb2Args, const double *, double *, double *, long long,
                                                      const double2 *, int, double *, int);
template __global__ void k_orb2v<false, true, false>(Orb2Args, const double *, double *, double *, long long,
                                                     const double2 *, int, double *, int);
template __global__ void k_orb2v<true, false, false>(Orb2Args, const double *, double *, double *, long long,
                                                     const double2 *, int, double *, int);
template __global__ void k_orb2v<true, false, true>(Orb2Args, const double *, double *, double *, long long,
                                                    const double2 *, int, double *, int);

// ------------------------------------------- vector-major two-group segments
template <bool BR, int G, int H>
__device__ __forceinline__ void orb2s_op(const Orb2Args &a, const unsigned long long *pw, int kk, uint32_t spect,
                                         double2 t, double (&x)[16], double (&l)[16], unsigned cmask, double &acc)
{
    const unsigned w = (unsigned)pw[kk] & 0xfffu;
    if (!w) return;
    const bool neg = __popc(spect & a.ops[a.k0 + kk].mask) & 1;
    double br = 0;
    orb2_apply_uniform<G, H>(w, neg, t.x, t.y, x, l, BR, br, (cmask >> (4 * G)) & 0xfu);
    if (BR) acc += br;
}

template <bool BR>
__global__ void __launch_bounds__(kBlock)
    k_orb2s(Orb2Args a, double *v0, double *v1, long long vstride, const double2 *__restrict__ cs, int M,
            double *__restrict__ partial, int slots)
{
    constexpr int K = kOrb2MaxOps;
    __shared__ double red[kBlock / 64];
    const int vb = blockIdx.y;
    v0 += vb * vstride;
    if (BR) v1 += vb * vstride;
    const int nk = a.k1 - a.k0, m = a.split;
    double2 t[K];
#pragma unroll
    for (int kk = 0; kk < K; kk++) t[kk] = kk < nk ? cs[(long long)vb * M + a.k0 + kk] : make_double2(1.0, 0.0);
    double acc[K];
#pragma unroll
    for (int kk = 0; kk < K; kk++) acc[kk] = 0;
    // orbit index space as in k_orb2v; the lone determinants (single-member
    // alpha and beta classes) are left alone: nothing acts on them
    const int mainA = a.nAm * a.nBc;
    const int total = mainA + (a.nAc - a.nAm) * a.nBm;
    for (int tt = blockIdx.x * kBlock + threadIdx.x; tt < total; tt += gridDim.x * kBlock) {
        int ia, ib;
        if (tt < mainA) {
            ia = tt / a.nBc;
            ib = tt - ia * a.nBc;
        } else {
            const int u = tt - mainA, i = u / a.nBm;
            ia = a.nAm + i;
            ib = u - i * a.nBm;
        }
        const int4 ra = a.aRanks[ia], rb = a.bRanks[ib];
        const uint32_t fa = a.aInfo[ia], fb = a.bInfo[ib];
        const int rA[4] = {ra.x, ra.y, ra.z, ra.w}, rB[4] = {rb.x, rb.y, rb.z, rb.w};
        const int na_ = (ra.x >= 0) + (ra.y >= 0) + (ra.z >= 0) + (ra.w >= 0);
        const int nb_ = (rb.x >= 0) + (rb.y >= 0) + (rb.z >= 0) + (rb.w >= 0);
        if (na_ * nb_ < 2) continue;
        const int type = ((fa >> 24) & 3) + 3 * ((fa >> 26) & 3) + 9 * (((fb >> 24) & 3) + 3 * ((fb >> 26) & 3));
        const uint32_t spect = ((fa & 0xffffffu) << a.n) | (fb & 0xffffffu);
        unsigned cmask;
        {
            const bool a2 = ((fa >> 26) & 3) == 0, b2 = ((fb >> 26) & 3) == 0;
            const bool a1 = ((fa >> 24) & 3) == 0, b1 = ((fb >> 24) & 3) == 0;
            cmask = 1u | (a2 ? 2u : 0u) | (b2 ? 4u : 0u) | (a2 && b2 ? 8u : 0u) |
                    ((1u | (a1 ? 2u : 0u) | (b1 ? 4u : 0u) | (a1 && b1 ? 8u : 0u)) << 4);
        }
        double x[16], l[16];
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const int ma = q & 3, mb = q >> 2;
            const bool ok = rA[ma] >= 0 && rB[mb] >= 0;
            const long long off = (long long)rA[ma] * a.nb + rB[mb];
            x[q] = ok ? v0[off] : 0.0;
            l[q] = (BR && ok) ? v1[off] : 0.0;
        }
        const unsigned long long *pw = a.prog + type * K;
        if (!BR) {
#pragma unroll
            for (int kk = 0; kk < K; kk++) {
                if (kk >= m) break;
                orb2s_op<BR, 0, 0>(a, pw, kk, spect, t[kk], x, l, cmask, acc[kk]);
            }
#pragma unroll
            for (int kk = 0; kk < K; kk++) {
                if (kk < m || kk >= nk) continue;
                orb2s_op<BR, 1, 1>(a, pw, kk, spect, t[kk], x, l, cmask, acc[kk]);
            }
        } else {
#pragma unroll
            for (int j = 0; j < K; j++) {
                const int kk = K - 1 - j;
                if (kk < m || kk >= nk) continue;
                orb2s_op<BR, 1, 1>(a, pw, kk, spect, t[kk], x, l, cmask, acc[kk]);
            }
#pragma unroll
            for (int j = 0; j < K; j++) {
                const int kk = K - 1 - j;
                if (kk >= m) continue;
                orb2s_op<BR, 0, 0>(a, pw, kk, spect, t[kk], x, l, cmask, acc[kk]);
            }
        }
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const int ma = q & 3, mb = q >> 2;
            if (rA[ma] >= 0 && rB[mb] >= 0) {
                const long long off = (long long)rA[ma] * a.nb + rB[mb];
                v0[off] = x[q];
                if (BR) v1[off] = l[q];
            }
        }
    }
    if (BR) {
#pragma unroll
        for (int kk = 0; kk < K; kk++) {
            if (kk >= nk) break;
            const double tot = block_sum(acc[kk], red);
            if (threadIdx.x == 0) partial[((long long)vb * M + a.k0 + kk) * slots + blockIdx.x] = tot;
        }
    }
}

template __global__ void k_orb2s<false>(Orb2Args, double *, double *, long long, const double2 *, int, double *, int);
template __global__ void k_orb2s<true>(Orb2Args, double *, double *, long long, const double2 *, int, double *, int);

// ------------------------------------------- Hessian tangents, two groups
// One operator of a tangent step on its active pairs in the sub-orbits the
// orbit type has (cm): psi, lambda and the lane's tangent rotate, the
// injection lane adds sc G psi, the brackets accumulate.
template <int G, int H>
__device__ __forceinline__ void orb2t_apply(unsigned w, bool neg, double C, double S, double (&ps)[16],
                                            double (&lm)[16], double (&b)[16], bool inj, double sc, unsigned cm,
                                            double &xacc, double &dacc)
{
    constexpr int PP[6] = {0, 0, 0, 1, 1, 2}, QQ[6] = {1, 2, 3, 2, 3, 3};
#pragma unroll
    for (int r = 0; r < 6; r++) {
        if (!((w >> r) & 1u)) continue;
        const double s = ((((w >> (6 + r)) & 1u) != 0) != neg) ? -1.0 : 1.0;
        const double sS = s * S;
        const int ip = PP[r] >> 1, jp = PP[r] & 1, iq = QQ[r] >> 1, jq = QQ[r] & 1;
#pragma unroll
        for (int c = 0; c < 4; c++) {
            if (c > 0 && !((cm >> c) & 1u)) continue;
            const int oa = c & 1, ob = c >> 1;
            const int sp = ((ip << G) | (oa << (1 - G))) + 4 * ((jp << H) | (ob << (1 - H)));
            const int sq = ((iq << G) | (oa << (1 - G))) + 4 * ((jq << H) | (ob << (1 - H)));
            const double pp = ps[sp], pq = ps[sq], lp = lm[sp], lq = lm[sq], bp = b[sp], bq = b[sq];
            const double pp2 = C * pp + sS * pq, pq2 = C * pq - sS * pp;
            const double lp2 = C * lp + sS * lq, lq2 = C * lq - sS * lp;
            double bp2 = C * bp + sS * bq, bq2 = C * bq - sS * bp;
            if (inj) {
                bp2 += sc * s * pq2;
                bq2 -= sc * s * pp2;
            }
            ps[sp] = pp2;
            ps[sq] = pq2;
            lm[sp] = lp2;
            lm[sq] = lq2;
            b[sp] = bp2;
            b[sq] = bq2;
            xacc += s * (lp2 * bq2 - lq2 * bp2);
            dacc -= lp2 * pp2 + lq2 * pq2;
        }
    }
}

template <int G, int H>
__device__ __forceinline__ void orb2t_op(const Orb2Args &a, const TanArgs &t, const unsigned long long *pw, int kk,
                                         uint32_t spect, const double2 (*csT)[1], int c, int lane, double (&ps)[16],
                                         double (&lm)[16], double (&b)[16], unsigned cmask, double *accW, double *dW)
{
    const unsigned w = (unsigned)pw[kk] & 0xfffu;
    if (!w) return;
    const int k = a.k0 + kk;
    const bool neg = __popc(spect & a.ops[k].mask) & 1;
    const double2 cs = csT[kk][0];
    const bool inj = t.inj[k] == c * kLanes + lane;
    double xacc = 0, dacc = 0;
    orb2t_apply<G, H>(w, neg, cs.x, cs.y, ps, lm, b, inj, t.scale[k], (cmask >> (4 * G)) & 0xfu, xacc, dacc);
    accW[kk * kLanes + lane] += xacc;
    if (lane == 0) dW[kk] += dacc;
}

__global__ void __launch_bounds__(kBlock, 4)
    k_orb2t(Orb2Args a, TanArgs t, double *B, long long cstride, const double2 *__restrict__ cs, int M)
{
    constexpr int WPB = kBlock / kLanes;
    constexpr int K = kOrb2MaxOps;
    __shared__ double2 csT[K][1];
    __shared__ double accL[WPB * K * kLanes];
    __shared__ double dL[WPB * K];
    const int lane = threadIdx.x & (kLanes - 1);
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / kLanes);
    const int c = blockIdx.y;
    const int nk = a.k1 - a.k0, m = a.split;
    if ((int)threadIdx.x < nk) csT[threadIdx.x][0] = cs[a.k0 + threadIdx.x];
    for (int e = threadIdx.x; e < WPB * K * kLanes; e += kBlock) accL[e] = 0.0;
    if ((int)threadIdx.x < WPB * K) dL[threadIdx.x] = 0.0;
    __syncthreads();
    double *accW = accL + wv * K * kLanes;
    double *dW = dL + wv * K;
    B += c * cstride;
    const int mainA = a.nAm * a.nBc;
    const int total = mainA + (a.nAc - a.nAm) * a.nBm;
    const int stride = gridDim.x * WPB;
    for (int tt = blockIdx.x * WPB + wv; tt < total; tt += stride) {
        int ia, ib;
        if (tt < mainA) {
            ia = tt / a.nBc;
            ib = tt - ia * a.nBc;
        } else {
            const int u = tt - mainA, i = u / a.nBm;
            ia = a.nAm + i;
            ib = u - i * a.nBm;
        }
        const int4 ra = ld_const(a.aRanks, ia), rb = ld_const(a.bRanks, ib);
        const uint32_t fa = ld_const(a.aInfo, ia), fb = ld_const(a.bInfo, ib);
        const int rA[4] = {ra.x, ra.y, ra.z, ra.w}, rB[4] = {rb.x, rb.y, rb.z, rb.w};
        const int na_ = (ra.x >= 0) + (ra.y >= 0) + (ra.z >= 0) + (ra.w >= 0);
        const int nb_ = (rb.x >= 0) + (rb.y >= 0) + (rb.z >= 0) + (rb.w >= 0);
        if (na_ * nb_ < 2) continue;  // a lone determinant: no operator acts
        const int type = ((fa >> 24) & 3) + 3 * ((fa >> 26) & 3) + 9 * (((fb >> 24) & 3) + 3 * ((fb >> 26) & 3));
        const uint32_t spect = ((fa & 0xffffffu) << a.n) | (fb & 0xffffffu);
        unsigned cmask;
        {
            const bool a2 = ((fa >> 26) & 3) == 0, b2 = ((fb >> 26) & 3) == 0;
            const bool a1 = ((fa >> 24) & 3) == 0, b1 = ((fb >> 24) & 3) == 0;
            cmask = 1u | (a2 ? 2u : 0u) | (b2 ? 4u : 0u) | (a2 && b2 ? 8u : 0u) |
                    ((1u | (a1 ? 2u : 0u) | (b1 ? 4u : 0u) | (a1 && b1 ? 8u : 0u)) << 4);
        }
        double ps[16], lm[16], b[16];
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const int ma = q & 3, mb = q >> 2;
            const bool ok = rA[ma] >= 0 && rB[mb] >= 0;
            const uint32_t d = (uint32_t)rA[ma] * (uint32_t)a.nb + (uint32_t)rB[mb];
            ps[q] = ok ? t.psi[d] : 0.0;
            lm[q] = ok ? t.lam[d] : 0.0;
            b[q] = ok ? B[(size_t)d * kLanes + lane] : 0.0;
        }
        const unsigned long long *pw = a.prog + type * K;
#pragma unroll
        for (int kk = 0; kk < K; kk++) {
            if (kk >= m) break;
            orb2t_op<0, 0>(a, t, pw, kk, spect, csT, c, lane, ps, lm, b, cmask, accW, dW);
        }
#pragma unroll
        for (int kk = 0; kk < K; kk++) {
            if (kk < m || kk >= nk) continue;
            orb2t_op<1, 1>(a, t, pw, kk, spect, csT, c, lane, ps, lm, b, cmask, accW, dW);
        }
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const int ma = q & 3, mb = q >> 2;
            if (rA[ma] >= 0 && rB[mb] >= 0)
                B[(size_t)((uint32_t)rA[ma] * (uint32_t)a.nb + (uint32_t)rB[mb]) * kLanes + lane] = b[q];
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < nk * kLanes; e += kBlock) {
        const int kk = e / kLanes, l = e - kk * kLanes;
        double s = 0;
#pragma unroll
        for (int q = 0; q < WPB; q++) s += accL[(q * K + kk) * kLanes + l];
        t.xpart[(((long long)c * M + a.k0 + kk) * t.slots + blockIdx.x) * kLanes + l] = s;
    }
    if (c == 0 && (int)threadIdx.x < nk) {
        double s = 0;
#pragma unroll
        for (int q = 0; q < WPB; q++) s += dL[q * K + threadIdx.x];
        t.dpart[(long long)(a.k0 + threadIdx.x) * t.slots + blockIdx.x] = s;
    }
}

// ------------------------------------------------ three-group segments
// f(Idx<I>) for I = B .. E-1, written out at compile time (no loop)
template <int I>
struct Idx {
    static constexpr int value = I;
};
template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F &&f)
{
    if constexpr (B < E) {
        f(Idx<B>{});
        static_for<B + 1, E>(f);
    }
}

// The value of the lane (this lane xor L) of its quad (L = 1, 2, 3): two
// 32-bit DPP quad_perm moves (DPP on 64-bit ALU operations takes only row
// broadcasts).
template <int L>
__device__ __forceinline__ double quad_xor(double v)
{
    constexpr int ctrl = L == 1 ? 0xb1 : (L == 2 ? 0x4e : 0x1b);
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), ctrl, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), ctrl, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}

// sum over the Q lanes of a vector (lanes Q v .. Q v + Q - 1)
template <int Q>
__device__ __forceinline__ double quad_sum(double v)
{
    if constexpr (Q >= 2) v += quad_xor<1>(v);
    if constexpr (Q == 4) v += quad_xor<2>(v);
    return v;
}

// bracket of one operator for this item: summed over the lanes of a vector,
// added by the first of them to the wave's row of the vector
template <int Q>
__device__ __forceinline__ void orb3_bracket(double *accW, int kk, int vv, int sl, double br)
{
    br = quad_sum<Q>(br);
    if (sl == 0) accW[kk * kLanes + vv] += br;
}

// operator of run 0 or 1 (groups 1, 2 in registers, as in k_orb2v)
template <bool BR, int G, int Q>
__device__ __forceinline__ void orb3_op12(const Orb3Args &a, const unsigned long long *pw, int kk, uint32_t spect,
                                          const double2 (*csL)[kLanes], double *accW, int vv, int sl,
                                          double (&x)[16], double (&l)[16], unsigned cmask)
{
    const unsigned w = (unsigned)pw[kk] & 0xfffu;
    if (!w) return;
    const bool neg = __popc(spect & a.ops[a.k0 + kk].mask) & 1;
    const double2 t = csL[kk][vv];
    double br = 0;
    orb2_apply_uniform<G, G>(w, neg, t.x, t.y, x, l, BR, br, BR ? (cmask >> (4 * G)) & 0xfu : 0xfu);
    if (BR) orb3_bracket<Q>(accW, kk, vv, sl, br);
}

// x <- cd x + cf xq, in place (as rot_inplace: a conditional update must
// leave the member arrays in their registers)
__device__ __forceinline__ void rot3_inplace(double &x, double xq, double cd, double cf)
{
    double t;
    // operands: %0 x, %1 temporary, %2 xq, %3 cf, %4 cd
    asm("v_mul_f64 %1, %3, %2\n\t"
        "v_fma_f64 %0, %4, %0, %1"
        : "+v"(x), "=&v"(t)
        : "v"(xq), "v"(cf), "v"(cd));
}

// adjoint: brr += pm (l xq - lq x), then x <- cd x + cf xq, l <- cd l + cf lq,
// in place
__device__ __forceinline__ void rot3_br_inplace(double &x, double &l, double &brr, double xq, double lq, double cd,
                                                double cf, double pm)
{
    double t1, t2, t3;
    // operands: %0 x, %1 l, %2 brr, %3-%5 temporaries, %6 xq, %7 lq, %8 cd,
    // %9 cf, %10 pm
    asm("v_mul_f64 %3, %1, %6\n\t"
        "v_fma_f64 %3, -%7, %0, %3\n\t"
        "v_fma_f64 %2, %10, %3, %2\n\t"
        "v_mul_f64 %4, %9, %6\n\t"
        "v_mul_f64 %5, %9, %7\n\t"
        "v_fma_f64 %0, %8, %0, %4\n\t"
        "v_fma_f64 %1, %8, %1, %5"
        : "+v"(x), "+v"(l), "+v"(brr), "=&v"(t1), "=&v"(t2), "=&v"(t3)
        : "v"(xq), "v"(lq), "v"(cd), "v"(cf), "v"(pm));
}

// operator of run 2: pairs of group-3 slots, one slot per lane; partner
// lane = this lane xor d.  The branch on d only picks the partner moves:
// around the in-place updates it made the compiler copy all 32 members at
// the merge.
template <bool BR>
__device__ __forceinline__ void orb3_rot3(int d, double cd, double cf, double pm, double (&x)[16],
                                          double (&l)[16], unsigned ex, double &brr)
{
#pragma unroll
    for (int j = 0; j < 16; j++) {
        if (BR && !((ex >> j) & 1u)) continue;
        double xq, lq = 0;
        if (d == 1) {
            xq = quad_xor<1>(x[j]);
            if (BR) lq = quad_xor<1>(l[j]);
        } else if (d == 2) {
            xq = quad_xor<2>(x[j]);
            if (BR) lq = quad_xor<2>(l[j]);
        } else {
            xq = quad_xor<3>(x[j]);
            if (BR) lq = quad_xor<3>(l[j]);
        }
        if (BR) rot3_br_inplace(x[j], l[j], brr, xq, lq, cd, cf, pm);
        else rot3_inplace(x[j], xq, cd, cf);
    }
}

template <bool BR, int QA, int QB>
__device__ __forceinline__ void orb3_op3(const Orb3Args &a, const unsigned long long *pw3, int kk, uint32_t spect,
                                         const double2 (*csL)[kLanes], double *accW, int vv, int sl, int slot,
                                         double (&x)[16], double (&l)[16], unsigned ex)
{
    constexpr int Q = QA * QB;
    constexpr int PP[6] = {0, 0, 0, 1, 1, 2}, QQ[6] = {1, 2, 3, 2, 3, 3};
    const unsigned w = (unsigned)pw3[kk] & 0xfffu;
    if (!w) return;
    const bool neg = __popc(spect & a.ops[a.k0 + kk].mask) & 1;
    const double2 t = csL[kk][vv];
    // this lane's pair: (p, q) <- (C p + sS q, C q - sS p); bracket s (l_p x_q - l_q x_p) on the p side
    double cd = 1.0, cf = 0.0, pm = 0.0;
    int d = 0;
#pragma unroll
    for (int r = 0; r < 6; r++) {
        if (!((w >> r) & 1u)) continue;
        d = PP[r] ^ QQ[r];
        const double s = ((((w >> (6 + r)) & 1u) != 0) != neg) ? -1.0 : 1.0;
        if (slot == PP[r]) {
            cd = t.x;
            cf = s * t.y;
            pm = s;
        } else if (slot == QQ[r]) {
            cd = t.x;
            cf = -s * t.y;
        }
    }
    double brr = 0;
    // slot bits (a3, b3) are lane bits 0, 1 with both spread, lane bit 0 with one
    orb3_rot3<BR>(Q == 4 ? d : 1, cd, cf, pm, x, l, ex, brr);
    if (BR) orb3_bracket<Q>(accW, kk, vv, sl, brr);
}

// one work item: orbit (ia, ib), vectors part * 64 / Q ..
template <bool BR, int QA, int QB>
__device__ __forceinline__ void orb3_item(const Orb3Args &a, int ia, int ib, int part, const double *v0, double *xout,
                                          double *v1, bool copyAll, const double2 (*csL)[kLanes], double *accW,
                                          int lane)
{
    constexpr int Q = QA * QB, VPI = kLanes / Q, K = kOrb3MaxOps;
    const int sl = lane & (Q - 1);
    const int a3 = QA == 2 ? (sl & 1) : 0;
    const int b3 = QB == 2 ? (QA == 2 ? (sl >> 1) : sl) : 0;
    const int slot = a3 + 2 * b3;
    const int vv = part * VPI + lane / Q;
    const int4 ra0 = ld_const(a.aRanks, 2 * ia), rb0 = ld_const(a.bRanks, 2 * ib);
    int rA[4] = {ra0.x, ra0.y, ra0.z, ra0.w}, rB[4] = {rb0.x, rb0.y, rb0.z, rb0.w};
    if constexpr (QA == 2) {
        const int4 r1 = ld_const(a.aRanks, 2 * ia + 1);
        rA[0] = a3 ? r1.x : rA[0];
        rA[1] = a3 ? r1.y : rA[1];
        rA[2] = a3 ? r1.z : rA[2];
        rA[3] = a3 ? r1.w : rA[3];
    }
    if constexpr (QB == 2) {
        const int4 r1 = ld_const(a.bRanks, 2 * ib + 1);
        rB[0] = b3 ? r1.x : rB[0];
        rB[1] = b3 ? r1.y : rB[1];
        rB[2] = b3 ? r1.z : rB[2];
        rB[3] = b3 ? r1.w : rB[3];
    }
    const uint32_t fa = ld_const(a.aInfo, ia), fb = ld_const(a.bInfo, ib);
    const int type = ((fa >> 24) & 3) + 3 * ((fa >> 26) & 3) + 9 * (((fb >> 24) & 3) + 3 * ((fb >> 26) & 3));
    const int type3 = ((fa >> 28) & 3) + 3 * ((fb >> 28) & 3);
    const uint32_t spect = ((fa & 0xffffffu) << a.n) | (fb & 0xffffffu);
    // members present per spin in groups 1, 2 (the same for every group-3 member)
    const bool a1 = ((fa >> 24) & 3) == 0, a2 = ((fa >> 26) & 3) == 0;
    const bool b1 = ((fb >> 24) & 3) == 0, b2 = ((fb >> 26) & 3) == 0;
    const unsigned exA = 1u | (a1 ? 2u : 0u) | (a2 ? 4u : 0u) | (a1 && a2 ? 8u : 0u);
    const unsigned exB = 1u | (b1 ? 2u : 0u) | (b2 ? 4u : 0u) | (b1 && b2 ? 8u : 0u);
    unsigned ex = 0;
#pragma unroll
    for (int mb = 0; mb < 4; mb++)
        if ((exB >> mb) & 1u) ex |= exA << (4 * mb);
    const unsigned cmask = 1u | (a2 ? 2u : 0u) | (b2 ? 4u : 0u) | (a2 && b2 ? 8u : 0u) |
                           ((1u | (a1 ? 2u : 0u) | (b1 ? 4u : 0u) | (a1 && b1 ? 8u : 0u)) << 4);
    // one member in every group: nothing acts on it
    const bool act = Q > 1 || __popc(ex) >= 2;
    const bool live = act || copyAll;
    if (!live) return;
    double x[16], l[16];
#pragma unroll
    for (int s = 0; s < 16; s++) {
        const int ma = s & 3, mb = s >> 2;
        const bool ok = rA[ma] >= 0 && rB[mb] >= 0;
        const uint32_t off = ((uint32_t)rA[ma] * (uint32_t)a.nb + (uint32_t)rB[mb]) * kLanes;
        x[s] = ok ? __builtin_nontemporal_load(v0 + off + vv) : 0.0;
        l[s] = (BR && ok) ? __builtin_nontemporal_load(v1 + off + vv) : 0.0;
    }
    if (act) {
        // runtime operator loops (the bodies are the same for every operator;
        // unrolled, the 48 bodies of the three runs overflowed the instruction
        // cache and the register allocator)
        const int nk = a.k1 - a.k0, s1 = a.s1, s2 = a.s2;
        const unsigned long long *pw = a.prog + type * K, *pw3 = a.prog3 + type3 * K;
        // every run (at most kOrbMaxOps operators) written out position by
        // position with a compile-time direction, as in k_orb2v: the member
        // arrays keep one register assignment throughout.  (A loop with the
        // partner moves, which are convergent, stayed a loop and copied all
        // 32 members at every iteration.)
        constexpr int R = kOrbMaxOps;
        if (!BR) {
            static_for<0, R>([&](auto I) {
                constexpr int i = decltype(I)::value;
                if (i < s1) orb3_op12<BR, 0, Q>(a, pw, i, spect, csL, accW, vv, sl, x, l, cmask);
            });
            static_for<0, R>([&](auto I) {
                constexpr int i = decltype(I)::value;
                if (s1 + i < s2) orb3_op12<BR, 1, Q>(a, pw, s1 + i, spect, csL, accW, vv, sl, x, l, cmask);
            });
            if constexpr (Q > 1)
                static_for<0, R>([&](auto I) {
                    constexpr int i = decltype(I)::value;
                    if (s2 + i < nk) orb3_op3<BR, QA, QB>(a, pw3, s2 + i, spect, csL, accW, vv, sl, slot, x, l, ex);
                });
        } else {
            if constexpr (Q > 1)
                static_for<0, R>([&](auto I) {
                    constexpr int i = decltype(I)::value;
                    if (nk - 1 - i >= s2)
                        orb3_op3<BR, QA, QB>(a, pw3, nk - 1 - i, spect, csL, accW, vv, sl, slot, x, l, ex);
                });
            static_for<0, R>([&](auto I) {
                constexpr int i = decltype(I)::value;
                if (s2 - 1 - i >= s1) orb3_op12<BR, 1, Q>(a, pw, s2 - 1 - i, spect, csL, accW, vv, sl, x, l, cmask);
            });
            static_for<0, R>([&](auto I) {
                constexpr int i = decltype(I)::value;
                if (s1 - 1 - i >= 0) orb3_op12<BR, 0, Q>(a, pw, s1 - 1 - i, spect, csL, accW, vv, sl, x, l, cmask);
            });
        }
    }
#pragma unroll
    for (int s = 0; s < 16; s++) {
        const int ma = s & 3, mb = s >> 2;
        if (rA[ma] >= 0 && rB[mb] >= 0) {
            const uint32_t off = ((uint32_t)rA[ma] * (uint32_t)a.nb + (uint32_t)rB[mb]) * kLanes;
            if (xout) __builtin_nontemporal_store(x[s], xout + off + vv);
            if (BR) __builtin_nontemporal_store(l[s], v1 + off + vv);
        }
    }
}

template <bool BR>
// forward: at most 128 VGPRs (4 waves per SIMD); adjoint: 168 (3 waves, as
// many as its 48 KB of LDS per block allow)
__global__ void __launch_bounds__(kBlock, BR ? 3 : 4)
    k_orb3v(Orb3Args a, const double *v0, double *xout, double *__restrict__ v1, long long cstride,
            const double2 *__restrict__ cs, int M, double *__restrict__ partial, int slots)
{
    constexpr int WPB = kBlock / kLanes;
    constexpr int K = kOrb3MaxOps;
    __shared__ double2 csL[K][kLanes];
    __shared__ double accL[BR ? WPB * K * kLanes : 1];
    const int lane = threadIdx.x & (kLanes - 1);
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / kLanes);
    const int c = blockIdx.y;
    const int nk = a.k1 - a.k0;
    for (int e = threadIdx.x; e < nk * kLanes; e += kBlock) {
        const int kk = e / kLanes, l = e - kk * kLanes;
        csL[kk][l] = cs[((long long)c * M + a.k0 + kk) * kLanes + l];
    }
    if (BR)
        for (int e = threadIdx.x; e < WPB * K * kLanes; e += kBlock) accL[e] = 0.0;
    __syncthreads();
    double *accW = BR ? accL + wv * K * kLanes : nullptr;
    v0 += c * cstride;
    if (xout) xout += c * cstride;
    if (BR) v1 += c * cstride;
    const bool copyAll = xout && xout != v0;
    const int nB1 = a.nB - a.nB2;
    const long long stride = (long long)gridDim.x * WPB;
    for (long long t = blockIdx.x * (long long)WPB + wv; t < a.w[3]; t += stride) {
        if (t < a.w[0]) {
            const int o = (int)(t >> 2), ia = o / a.nB2;
            orb3_item<BR, 2, 2>(a, ia, o - ia * a.nB2, (int)(t & 3), v0, xout, v1, copyAll, csL, accW, lane);
        } else if (t < a.w[1]) {
            const long long u = t - a.w[0];
            const int o = (int)(u >> 1), ia = o / nB1;
            orb3_item<BR, 2, 1>(a, ia, a.nB2 + o - ia * nB1, (int)(u & 1), v0, xout, v1, copyAll, csL, accW,
                                lane);
        } else if (t < a.w[2]) {
            const long long u = t - a.w[1];
            const int o = (int)(u >> 1), i = o / a.nB2;
            orb3_item<BR, 1, 2>(a, a.nA2 + i, o - i * a.nB2, (int)(u & 1), v0, xout, v1, copyAll, csL, accW,
                                lane);
        } else {
            const int o = (int)(t - a.w[2]), i = o / nB1;
            orb3_item<BR, 1, 1>(a, a.nA2 + i, a.nB2 + o - i * nB1, 0, v0, xout, v1, copyAll, csL, accW, lane);
        }
    }
    if (BR) {
        __syncthreads();
        for (int e = threadIdx.x; e < nk * kLanes; e += kBlock) {
            const int kk = e / kLanes, l = e - kk * kLanes;
            double s = 0;
#pragma unroll
            for (int q = 0; q < WPB; q++) s += accL[(q * K + kk) * kLanes + l];
            partial[(((long long)c * M + a.k0 + kk) * slots + blockIdx.x) * kLanes + l] = s;
        }
    }
}

template __global__ void k_orb3v<false>(Orb3Args, const double *, double *, double *, long long, const double2 *, int,
                                        double *, int);
template __global__ void k_orb3v<true>(Orb3Args, const double *, double *, double *, long long, const double2 *, int,
                                       double *, int);

template <bool BR>
__global__ void __launch_bounds__(kBlock)
    k_opv(double *__restrict__ v0, double *__restrict__ v1, long long cstride, long long N, DevSector sec,
          const DevOp *__restrict__ ops, int k, const double2 *__restrict__ cs, int M, double *__restrict__ partial,
          int slots)
{
    constexpr int WPB = kBlock / kLanes;
    __shared__ double red[BR ? WPB * kLanes : 1];
    const int lane = threadIdx.x & (kLanes - 1);
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / kLanes);
    const int c = blockIdx.y;
    v0 += c * cstride + lane;
    if (BR) v1 += c * cstride + lane;
    const DevOp o = ops[k];
    const double2 t = cs[((long long)c * M + k) * kLanes + lane];
    const double C = t.x, S = t.y;
    double acc[1] = {0};
    for (long long i = blockIdx.x * WPB + wv; i < N; i += (long long)gridDim.x * WPB) {
        uint32_t b2;
        const uint32_t b = sec.states[i];
        const int s = pair_sign(o, b, b2);
        if (s == 0 || b2 < b) continue;
        const long long j = sec.rank(b2);
        if (j < 0) continue;
        const double x = v0[i * kLanes], y = v0[j * kLanes];
        if (BR) {
            const double lx = v1[i * kLanes], ly = v1[j * kLanes];
            acc[0] += s * (lx * y - ly * x);
            v1[i * kLanes] = s * ly * S + lx * C;
            v1[j * kLanes] = -s * lx * S + ly * C;
        }
        v0[i * kLanes] = s * y * S + x * C;
        v0[j * kLanes] = -s * x * S + y * C;
    }
    if (BR) lanes_block_partials<1>(acc, 1, k, k + 1, 0, red, partial, c, M, slots);
}

template __global__ void k_orbv<false>(OrbArgs, const double *, double *, double *, long long, const double2 *, int,
                                       double *, int);
template __global__ void k_orbv<true>(OrbArgs, const double *, double *, double *, long long, const double2 *, int,
                                      double *, int);
template __global__ void k_opv<false>(double *, double *, long long, long long, DevSector, const DevOp *, int,
                                      const double2 *, int, double *, int);
template __global__ void k_opv<true>(double *, double *, long long, long long, DevSector, const DevOp *, int,
                                     const double2 *, int, double *, int);

} // namespace ae
