// Window passes of the lane-layout sweeps (k_win): a tile of determinants
// that agree outside a window of spatial orbitals is staged in LDS for the 64
// vectors of a chunk (lane = vector), every phase (orbit segment) of the pass
// is applied to it there, and it is written back once.  Every index, class
// and sign is wave-uniform and comes through the scalar unit; the vector
// unit does the rotations.  See kernels.hpp.
#include "kernels.hpp"

namespace ae {

// Read-only tables indexed by wave-uniform values, through the constant
// address space (scalar loads).
__device__ __forceinline__ uint32_t wld(const uint32_t *p, int i)
{
    return ((const __attribute__((address_space(4))) uint32_t *)p)[i];
}
__device__ __forceinline__ int wldi(const uint32_t *p, int i) { return (int)wld(p, i); }

// One phase of a pass on the tile in LDS: each wave takes sub-orbits
// (alpha sub-class, beta sub-class with at least one two-member side, up to
// 2 x 2 members) and rotates them in registers by the phase's orbit program,
// as one orbit of k_orbv.  acc[j]: bracket of operator nk - 1 - j (adjoint).
template <bool BR, int NW>
__device__ __forceinline__ void win_phase(const WinArgs &a, int p, int ca, int cb, int nB, uint32_t aOut,
                                          uint32_t bOut, int cOffA, int cOffB, double *X, double *L,
                                          const double2 *__restrict__ cs, int c, int M, double (&acc)[kOrbMaxOps])
{
    constexpr int PP[6] = {0, 0, 0, 1, 1, 2}, QQ[6] = {1, 2, 3, 2, 3, 3};
    const int lane = threadIdx.x & (kLanes - 1);
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / kLanes);
    const uint32_t *B = a.blob;
    const int k0 = wldi(B, p * 8 + 0), nk = wldi(B, p * 8 + 1), progOff = wldi(B, p * 8 + 2),
              kp = wldi(B, p * 8 + 3);
    const uint32_t spectM = wld(B, p * 8 + 6);
    const int ia4 = a.oSubIdx + ((p * 2) * (a.w + 1) + ca) * 4;
    const int ib4 = a.oSubIdx + ((p * 2 + 1) * (a.w + 1) + cb) * 4;
    const int offA = wldi(B, ia4), nA_ = wldi(B, ia4 + 1);
    const int offB = wldi(B, ib4), nB_ = wldi(B, ib4 + 1);
    const uint32_t magB = wld(B, ib4 + 2), magB2 = wld(B, ib4 + 3);
    const int naA = nA_ & 0xffff, n2A = nA_ >> 16, nbB = nB_ & 0xffff, n2B = nB_ >> 16;
    const int mainA = n2A * nbB;  // two-member alpha sub-classes x every beta sub-class
    const int work = mainA + (naA - n2A) * n2B;
    if (wv >= work) return;
    // the phase's angles for this lane's vector, and its sign masks
    double C[kOrbMaxOps], S[kOrbMaxOps];
    uint32_t msk[kOrbMaxOps];
#pragma unroll
    for (int j = 0; j < kOrbMaxOps; j++) {
        if (j >= nk) break;
        const double2 t = cs[((long long)c * M + k0 + j) * kLanes + lane];
        C[j] = t.x;
        S[j] = t.y;
        msk[j] = wld(B, a.oMask + kp + j);
    }
    for (int so = wv; so < work; so += NW) {
        uint32_t ea, eb;
        if (so < mainA) {
            ea = (uint32_t)(((uint64_t)so * magB) >> 24);
            eb = so - ea * nbB;
        } else {
            const uint32_t u = so - mainA, i = (uint32_t)(((uint64_t)u * magB2) >> 24);
            ea = n2A + i;
            eb = u - i * n2B;
        }
        const uint32_t EA = wld(B, a.oSub + offA + ea), EB = wld(B, a.oSub + offB + eb);
        const int shA = (int)(EA >> 16), shB = (int)(EB >> 16);
        const int ia[2] = {(int)(EA & 0xffu), (int)((EA >> 8) & 0xffu)};
        const int ib[2] = {(int)(EB & 0xffu), (int)((EB >> 8) & 0xffu)};
        const int Ra = shA == 0 ? 2 : 1, Rb = shB == 0 ? 2 : 1;
        const uint32_t spect =
            (((aOut | wld(B, a.oPat + cOffA + ia[0])) << a.n) | (bOut | wld(B, a.oPat + cOffB + ib[0]))) & spectM;
        const int prog = a.oProg + progOff + (shA * 3 + shB) * kOrbMaxOps;
        int m[4];
        bool ok[4];
        double x[4], l[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int i = q >> 1, j = q & 1;
            ok[q] = i < Ra && j < Rb;
            m[q] = (ok[q] ? ia[i] * nB + ib[j] : 0) * kLanes + lane;
            x[q] = ok[q] ? X[m[q]] : 0.0;
            l[q] = (BR && ok[q]) ? L[m[q]] : 0.0;
        }
#pragma unroll
        for (int j = 0; j < kOrbMaxOps; j++) {
            if (j >= nk) break;
            const int kk = BR ? nk - 1 - j : j;
            const uint32_t w = wld(B, prog + kk);
            if (!(w & 0x3fu)) continue;
            const bool neg = __popc(spect & msk[kk]) & 1;
            const double Cc = C[kk], Ss = S[kk];
            double br = 0;
#pragma unroll
            for (int r = 0; r < 6; r++) {
                if (!((w >> r) & 1u)) continue;
                const int pp = PP[r], qq = QQ[r];
                const double s = ((((w >> (8 + r)) & 1u) != 0) != neg) ? -1.0 : 1.0;
                const double xp = x[pp], xq = x[qq];
                if (BR) {
                    const double lp = l[pp], lq = l[qq];
                    br += s * (lp * xq - lq * xp);
                    l[pp] = s * lq * Ss + lp * Cc;
                    l[qq] = -s * lp * Ss + lq * Cc;
                }
                x[pp] = s * xq * Ss + xp * Cc;
                x[qq] = -s * xp * Ss + xq * Cc;
            }
            if (BR) acc[j] += br;
        }
#pragma unroll
        for (int q = 0; q < 4; q++)
            if (ok[q]) {
                X[m[q]] = x[q];
                if (BR) L[m[q]] = l[q];
            }
    }
}

template <bool BR, bool BC, int NW>
__global__ void __launch_bounds__(NW * kLanes)
    k_win(WinArgs a, const double *in, double *out, double *lam, long long cstride, const double2 *__restrict__ cs,
          int M, double *__restrict__ partial, int slots)
{
    extern __shared__ __align__(16) double lds[];
    const int lane = threadIdx.x & (kLanes - 1);
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / kLanes);
    const int c = blockIdx.y;
    double *X = lds;
    double *L = X + (size_t)a.maxMembers * kLanes;
    double *accP = L + (BR ? (size_t)a.maxMembers * kLanes : 0);  // per wave, one phase
    double *brk = accP + (BR ? NW * kOrbMaxOps * kLanes : 0);     // the pass's operators
    const uint32_t *B = a.blob;
    if (BR)
        for (int e = threadIdx.x; e < a.nkp * kLanes; e += NW * kLanes) brk[e] = 0.0;
    const long long base = (long long)c * cstride + lane;
    for (int it = blockIdx.x; it < a.ntiles; it += gridDim.x) {
        const int tile = __builtin_amdgcn_readfirstlane(a.tiles[it]);
        const int ja = tile / a.nBcls, jb = tile - ja * a.nBcls;
        const uint32_t fa = wld(B, a.oCls + ja), fb = wld(B, a.oCls + a.nAcls + jb);
        const uint32_t aOut = fa & 0xffffffu, bOut = fb & 0xffffffu;
        const int ca = (int)(fa >> 24), cb = (int)(fb >> 24);
        const int cOffA = wldi(B, a.oCOff + ca), nA = wldi(B, a.oCOff + ca + 1) - cOffA;
        const int cOffB = wldi(B, a.oCOff + cb), nB = wldi(B, a.oCOff + cb + 1) - cOffB;
        const int sA = wldi(B, a.oCStart + ja), sB = a.nrowsA + wldi(B, a.oCStart + a.nAcls + jb);
        const int nm = nA * nB;
        __syncthreads();  // the previous tile's stores have read X / L
        for (int mm = wv; mm < nm; mm += NW) {
            const int i = mm / nB, j = mm - i * nB;
            const long long g = (long long)wld(a.rows, sA + i) + wld(a.rows, sB + j);
            X[mm * kLanes + lane] = BC ? in[g] : __builtin_nontemporal_load(in + base + g * kLanes);
            if (BR) L[mm * kLanes + lane] = __builtin_nontemporal_load(lam + base + g * kLanes);
        }
        __syncthreads();
        for (int pi = 0; pi < a.nph; pi++) {
            const int p = BR ? a.nph - 1 - pi : pi;
            double acc[kOrbMaxOps];
#pragma unroll
            for (int kk = 0; kk < kOrbMaxOps; kk++) acc[kk] = 0.0;
            win_phase<BR, NW>(a, p, ca, cb, nB, aOut, bOut, cOffA, cOffB, X, L, cs, c, M, acc);
            if (BR) {
                const int nk = wldi(B, p * 8 + 1), kp = wldi(B, p * 8 + 3);
                __syncthreads();  // every wave has folded the previous phase's rows
#pragma unroll
                for (int j = 0; j < kOrbMaxOps; j++) {
                    if (j >= nk) break;
                    accP[(wv * kOrbMaxOps + (nk - 1 - j)) * kLanes + lane] = acc[j];
                }
                __syncthreads();
                // fixed-order sum over the waves; wave w folds operators w, w + NW, ...
                for (int kk = wv; kk < nk; kk += NW) {
                    double s = 0;
#pragma unroll
                    for (int q = 0; q < NW; q++) s += accP[(q * kOrbMaxOps + kk) * kLanes + lane];
                    brk[(kp + kk) * kLanes + lane] += s;
                }
            } else {
                __syncthreads();
            }
        }
        for (int mm = wv; mm < nm; mm += NW) {
            const int i = mm / nB, j = mm - i * nB;
            const long long g = (long long)wld(a.rows, sA + i) + wld(a.rows, sB + j);
            if (out) __builtin_nontemporal_store(X[mm * kLanes + lane], out + base + g * kLanes);
            if (BR) __builtin_nontemporal_store(L[mm * kLanes + lane], lam + base + g * kLanes);
        }
    }
    if (BR) {
        __syncthreads();
        for (int kp = wv; kp < a.nkp; kp += NW)
            partial[(((long long)c * M + a.kpOp[kp]) * slots + blockIdx.x) * kLanes + lane] = brk[kp * kLanes + lane];
    }
}

#define AE_WIN(BR, BC, NW)                                                                                  \
    template __global__ void k_win<BR, BC, NW>(WinArgs, const double *, double *, double *, long long,     \
                                               const double2 *__restrict__, int, double *__restrict__, int);
AE_WIN(false, false, 4)
AE_WIN(false, true, 4)
AE_WIN(true, false, 4)
AE_WIN(false, false, 8)
AE_WIN(false, true, 8)
AE_WIN(true, false, 8)
#undef AE_WIN

} // namespace ae

// ------------------------------------------- register windows (k_winr)
namespace ae {
namespace wr {

constexpr int C4[5] = {1, 4, 6, 4, 1};
constexpr int popc4(int x) { return (x & 1) + ((x >> 1) & 1) + ((x >> 2) & 1) + ((x >> 3) & 1); }
// the i-th 4-bit pattern with c bits set, ascending
constexpr int pat(int c, int i)
{
    int k = 0;
    for (int x = 0; x < 16; x++)
        if (popc4(x) == c) {
            if (k == i) return x;
            k++;
        }
    return 0;
}
constexpr int index_of(int x)
{
    int k = 0;
    for (int y = 0; y < x; y++)
        if (popc4(y) == popc4(x)) k++;
    return k;
}
// Layouts: a tile's registers hold window pattern x at index
// index_of(rot^r(x)), rot the cyclic right rotation of the four window bits;
// layout r = lp brings the pair (lp, lp + 1) to bits (0, 1), so one compiled
// form of the phases (pair 0) serves every pair.
constexpr int rot1(int x) { return ((x >> 1) | ((x & 1) << 3)) & 0xf; }
constexpr int rotk(int x, int k) { return k == 0 ? x : rotk(rot1(x), k - 1); }
// the register index of member i (count c) after K more rotations
constexpr int rperm(int c, int i, int K) { return index_of(rotk(pat(c, i), K)); }
// Sub-classes of the patterns with c bits w.r.t. the local pair (lp, lp + 1):
// patterns agreeing outside the pair; two members (one bit of the pair set:
// i0 the lower bit, i1 the upper) first, then single ones (shape 1: pair
// empty, 2: pair full).
struct Sub {
    int i0, i1, sh;
};
constexpr int nsub(int c, int lp)
{
    const int m = 3 << lp;
    int k = 0;
    for (int i = 0; i < C4[c]; i++) {
        const int x = pat(c, i);
        if ((x & m) != (1 << lp)) k++;  // one entry per two-member pair (its upper member), or a single
    }
    return k;
}
constexpr Sub sub(int c, int lp, int k)
{
    const int m = 3 << lp;
    int q = 0;
    for (int i = 0; i < C4[c]; i++) {  // two-member sub-classes
        const int x = pat(c, i);
        if ((x & m) == (1 << lp)) {
            if (q == k) return Sub{i, index_of(x ^ m), 0};
            q++;
        }
    }
    for (int i = 0; i < C4[c]; i++) {  // single members
        const int x = pat(c, i);
        if ((x & m) == 0 || (x & m) == m) {
            if (q == k) return Sub{i, -1, (x & m) ? 2 : 1};
            q++;
        }
    }
    return Sub{0, -1, 1};
}

template <int I>
struct Ix {
    static constexpr int value = I;
};
template <int B, int E, typename F>
__device__ __forceinline__ void wfor(F &&f)
{
    if constexpr (B < E) {
        f(Ix<B>{});
        wfor<B + 1, E>(f);
    }
}

__device__ __forceinline__ uint32_t ld(const uint32_t *p, int i)
{
    return ((const __attribute__((address_space(4))) uint32_t *)p)[i];
}

// parity of the bits of a 4-bit constant selected by m
template <int X>
__device__ __forceinline__ bool par4(uint32_t m)
{
    return __popc(m & (uint32_t)X) & 1;
}

// (p, q) <- (C p + sS q, C q - sS p), results in the registers of p and q
__device__ __forceinline__ void rot(double &p, double &q, double C, double sS)
{
    double t, u;
    asm("v_mul_f64 %2, %5, %1\n\t"
        "v_mul_f64 %3, %5, %0\n\t"
        "v_fma_f64 %0, %4, %0, %2\n\t"
        "v_fma_f64 %1, %4, %1, -%3"
        : "+v"(p), "+v"(q), "=&v"(t), "=&v"(u)
        : "v"(C), "v"(sS));
}
// adjoint step: brr += lp xq - lq xp (before the step), then both rotations
__device__ __forceinline__ void rot2(double &xp, double &xq, double &lp, double &lq, double &brr, double C,
                                     double sS)
{
    double t2, u2, t3, u3;
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

// x <- the registers relabeled by K more rotations (a compile-time permutation)
template <int CA, int CB, int K, int NM>
__device__ __forceinline__ void relabel(double (&x)[NM])
{
    constexpr int NB = C4[CB];
    double y[NM];
    wfor<0, NM>([&](auto I) {
        constexpr int m = decltype(I)::value;
        y[rperm(CA, m / NB, K) * NB + rperm(CB, m % NB, K)] = x[m];
    });
    wfor<0, NM>([&](auto I) { x[decltype(I)::value] = y[decltype(I)::value]; });
}
// TUPS units: operator kinds by position in the unit (0 beta single, 1
// alpha single, 2 paired double), the same in every unit; the member pairs
// an operator rotates then depend only on the orbit type, known at compile
// time, and the phases are straight-line code.  The host checks every
// program word against this pattern before it selects the form.
constexpr int kTupsKind[5] = {0, 1, 2, 0, 1};
// member-pair slots (0,1) (0,2) (0,3) (1,2) (1,3) (2,3) a kind rotates in an
// orbit of alpha shape sa and beta shape sb (0: two members)
constexpr unsigned tups_slots(int kind, int sa, int sb)
{
    return kind == 0 ? (sb == 0 ? (1u | (sa == 0 ? 32u : 0u)) : 0u)
         : kind == 1 ? (sa == 0 ? (2u | (sb == 0 ? 16u : 0u)) : 0u)
                     : (sa == 0 && sb == 0 ? 4u : 0u);
}

// d with its sign flipped when sg is 0x80000000 (sg 0: unchanged)
__device__ __forceinline__ double flip(double d, uint32_t sg)
{
    return __hiloint2double(__double2hiint(d) ^ (int)sg, __double2loint(d));
}

// One TUPS operator: its member pairs are compile-time and, its sign mask
// being empty, one phase serves every pair it rotates; the caller passes S
// with that phase applied.  The adjoint's bracket sum is returned without it.
template <int CA, int CB, int KIND, bool BR>
__device__ __forceinline__ void op_tups(double C, double S, double (&x)[C4[CA] * C4[CB]],
                                        double (&l)[BR ? C4[CA] * C4[CB] : 1], double &br)
{
    constexpr int NB = C4[CB];
    constexpr int NSA = nsub(CA, 0), NSB = nsub(CB, 0);
    constexpr int PP[6] = {0, 0, 0, 1, 1, 2}, QQ[6] = {1, 2, 3, 2, 3, 3};
    wfor<0, NSA * NSB>([&](auto I) {
        constexpr Sub A = sub(CA, 0, decltype(I)::value / NSB);
        constexpr Sub Bs = sub(CB, 0, decltype(I)::value % NSB);
        constexpr unsigned slots = tups_slots(KIND, A.sh, Bs.sh);
        if constexpr (slots != 0) {
            double brs = 0;
            wfor<0, 6>([&](auto R) {
                constexpr int r = decltype(R)::value;
                if constexpr ((slots >> r) & 1u) {
                    constexpr int p = PP[r], q = QQ[r];
                    constexpr int mp = ((p >> 1) ? A.i1 : A.i0) * NB + ((p & 1) ? Bs.i1 : Bs.i0);
                    constexpr int mq = ((q >> 1) ? A.i1 : A.i0) * NB + ((q & 1) ? Bs.i1 : Bs.i0);
                    if constexpr (BR) rot2(x[mp], x[mq], l[mp], l[mq], brs, C, S);
                    else rot(x[mp], x[mq], C, S);
                }
            });
            if constexpr (BR) br += brs;
        }
    });
}

template <int CA, int CB, bool BR, int NM, int NL>
__device__ __forceinline__ void to_layout(int d, double (&x)[NM], double (&l)[NL])
{
    if (d & 1) {
        relabel<CA, CB, 1>(x);
        if constexpr (BR) relabel<CA, CB, 1>(l);
    }
    if (d & 2) {
        relabel<CA, CB, 2>(x);
        if constexpr (BR) relabel<CA, CB, 2>(l);
    }
}
__device__ __forceinline__ uint32_t rot4(uint32_t m, int k)
{
    return ((m >> k) | (m << (4 - k))) & 0xfu;
}

// one operator's angles, tile sign, mask window bits and program words
struct wr_op_args {
    double2 cs;
    bool negT;
    uint32_t mA, mB;
    int progT;
};

struct TileCtx {
    uint32_t aOut, bOut;
    int sA, sB;  // first members in rows
};

// One operator of a phase on local pair LP over every sub-orbit of the tile.
// Its program words are read per orbit type.  The spectator sign of a
// sub-orbit is the tile's part (negT: outer bits) times the parity of its
// window bits under the operator's mask (mA, mB: the mask's window bits of
// each spin, with the spectator mask applied).
template <int CA, int CB, int LP, bool BR>
__device__ __forceinline__ void op_pair(const WinArgs &a, int progT, bool negT, uint32_t mA, uint32_t mB, double C,
                                        double S, double (&x)[C4[CA] * C4[CB]],
                                        double (&l)[BR ? C4[CA] * C4[CB] : 1], double &br)
{
    constexpr int NB = C4[CB];
    constexpr int NSA = nsub(CA, LP), NSB = nsub(CB, LP);
    constexpr int PP[6] = {0, 0, 0, 1, 1, 2}, QQ[6] = {1, 2, 3, 2, 3, 3};
    wfor<0, NSA * NSB>([&](auto I) {
        constexpr Sub A = sub(CA, LP, decltype(I)::value / NSB);
        constexpr Sub Bs = sub(CB, LP, decltype(I)::value % NSB);
        if constexpr (A.sh == 0 || Bs.sh == 0) {
            constexpr int type = A.sh * 3 + Bs.sh;
            const uint32_t w = ld(a.blob, progT + type);
            const bool neg = negT ^ par4<pat(CA, A.i0)>(mA) ^ par4<pat(CB, Bs.i0)>(mB);
            double brs = 0;
            wfor<0, 6>([&](auto R) {
                constexpr int r = decltype(R)::value;
                constexpr int p = PP[r], q = QQ[r];
                constexpr int ip = p >> 1, jp = p & 1, iq = q >> 1, jq = q & 1;
                constexpr bool okp = (ip == 0 || A.sh == 0) && (jp == 0 || Bs.sh == 0);
                constexpr bool okq = (iq == 0 || A.sh == 0) && (jq == 0 || Bs.sh == 0);
                if constexpr (okp && okq) {
                    constexpr int mp = (ip ? A.i1 : A.i0) * NB + (jp ? Bs.i1 : Bs.i0);
                    constexpr int mq = (iq ? A.i1 : A.i0) * NB + (jq ? Bs.i1 : Bs.i0);
                    if ((w >> r) & 1u) {
                        const bool minus = (((w >> (8 + r)) & 1u) != 0) != neg;
                        const double sS = minus ? -S : S;
                        if constexpr (BR) {
                            double brr = 0;
                            rot2(x[mp], x[mq], l[mp], l[mq], brr, C, sS);
                            brs += minus ? -brr : brr;
                        } else {
                            rot(x[mp], x[mq], C, sS);
                        }
                    }
                }
            });
            if constexpr (BR) br += brs;
        }
        // one sub-orbit at a time: interleaving them raises the live registers past the budget
        __builtin_amdgcn_sched_barrier(0);
    });
}

// The operators of one phase on the tile, in order (the adjoint in
// reverse), the tile in layout lp (its pair at bits 0, 1).  Angles: the
// forward reads them from LDS (csL, the pass's operators); the adjoint from
// the table, two operators ahead.
template <int CA, int CB, bool BR, bool TU>
__device__ __forceinline__ void phase_run(const WinArgs &a, const TileCtx &t, int p, int lp,
                                          const double2 *__restrict__ cs, const double2 *csL, int c, int M,
                                          double *brk, double (&x)[C4[CA] * C4[CB]],
                                          double (&l)[BR ? C4[CA] * C4[CB] : 1])
{
    const int lane = threadIdx.x & (kLanes - 1);
    const uint32_t *B = a.blob;
    const int k0 = (int)ld(B, p * 8 + 0), nk = (int)ld(B, p * 8 + 1), kp = (int)ld(B, p * 8 + 3);
    const uint32_t spectM = ld(B, p * 8 + 6);
    const uint32_t wm = 0xfu << a.w0;
    const uint32_t outer = ((t.aOut << a.n) | t.bOut) & ~((wm << a.n) | wm);
    const double2 *cp = cs + ((long long)c * M + k0) * kLanes + lane;
    if constexpr (TU) {
        // a TUPS unit: five operators, straight-line
        wfor<0, 5>([&](auto J) {
            constexpr int kk = BR ? 4 - decltype(J)::value : decltype(J)::value;
            const double2 cst = BR ? cp[kk * kLanes] : csL[(kp + kk) * kLanes + lane];
            const uint32_t sg = ld(B, a.oProgT + (p * kOrbMaxOps + kk) * 16 + 9);
            double br = 0;
            op_tups<CA, CB, kTupsKind[kk], BR>(cst.x, flip(cst.y, sg), x, l, br);
            if constexpr (BR) brk[(kp + kk) * kLanes + lane] += flip(br, sg);
        });
        return;
    }
    double2 c1 = make_double2(1.0, 0.0), c2 = c1;
    if constexpr (BR) {
        c1 = cp[(nk - 1) * kLanes];
        if (nk > 1) c2 = cp[(nk - 2) * kLanes];
    }
    for (int jj = 0; jj < nk; jj++) {
        const int kk = BR ? nk - 1 - jj : jj;
        double2 cst;
        if constexpr (BR) {
            cst = c1;
            c1 = c2;
            if (jj + 2 < nk) c2 = cp[(kk - 2) * kLanes];
        } else {
            cst = csL[(kp + kk) * kLanes + lane];
        }
        const uint32_t m = ld(B, a.oMask + kp + kk) & spectM;
        const bool negT = __popc(m & outer) & 1;
        // the mask's window bits in the tile's layout
        const uint32_t mA = rot4((m >> (a.n + a.w0)) & 0xfu, lp), mB = rot4((m >> a.w0) & 0xfu, lp);
        double br = 0;
        op_pair<CA, CB, 0, BR>(a, a.oProgT + (p * kOrbMaxOps + kk) * 16, negT, mA, mB, cst.x, cst.y, x, l, br);
        if constexpr (BR) brk[(kp + kk) * kLanes + lane] += br;
    }
}

// Member (i, j) of a tile sits at row rows[sA + i] (alpha rank * nb) plus
// column rows[sB + j] (beta rank).  Its address is a per-row base in scalar
// registers (NA of them) plus a per-column byte offset in a vector register
// (NB of them, lane included), so each load and store is one instruction on
// a scalar base and a vector offset, with no address arithmetic per member.
__device__ __forceinline__ double ld_at(const double *row, uint32_t boff)
{
    return __builtin_nontemporal_load((const double *)((const char *)row + boff));
}
__device__ __forceinline__ void st_at(double v, double *row, uint32_t boff)
{
    __builtin_nontemporal_store(v, (double *)((char *)row + boff));
}

// en (adjoint): also add the tile's <x|l> to e before the first phase
template <int CA, int CB, bool BR, bool BC, int NW, bool TU>
__device__ __forceinline__ void tile(const WinArgs &a, const TileCtx &t, const double *in, double *out, double *lam,
                                     long long base, const double2 *__restrict__ cs, const double2 *csL, int c, int M,
                                     double *brk, bool en, double &e)
{
    constexpr int NA = C4[CA], NB = C4[CB], NM = NA * NB;
    const uint32_t lane8 = (threadIdx.x & (kLanes - 1)) * 8u;
    double x[NM], l[BR ? NM : 1];
    {
        uint32_t col[NB], colS[NB];
        wfor<0, NB>([&](auto J) {
            constexpr int j = decltype(J)::value;
            const uint32_t rb = ld(a.rows, t.sB + j);
            col[j] = rb * (kLanes * 8u) + lane8;
            colS[j] = rb * 8u;  // the start vector: one amplitude per determinant
        });
        wfor<0, NA>([&](auto I) {
            constexpr int i = decltype(I)::value;
            const long long ra = ld(a.rows, t.sA + i);
            const double *rx = BC ? in + ra : in + base + ra * kLanes;
            const double *rl = lam + base + ra * kLanes;
            wfor<0, NB>([&](auto J) {
                constexpr int j = decltype(J)::value;
                if constexpr (BC) x[i * NB + j] = *(const double *)((const char *)rx + colS[j]);
                else x[i * NB + j] = ld_at(rx, col[j]);
                if constexpr (BR) l[i * NB + j] = ld_at(rl, col[j]);
            });
        });
    }
    if constexpr (BR)
        if (en) {
            double s = 0;
            wfor<0, NM>([&](auto I) { s = fma(x[decltype(I)::value], l[decltype(I)::value], s); });
            e += s;
        }
    int r = 0;  // the registers' layout
    for (int pi = 0; pi < a.nph; pi++) {
        const int p = BR ? a.nph - 1 - pi : pi;
        const int lp = __builtin_ctz(ld(a.blob, p * 8 + 4)) - a.w0;
        to_layout<CA, CB, BR>((lp - r) & 3, x, l);
        r = lp;
        phase_run<CA, CB, BR, TU>(a, t, p, lp, cs, csL, c, M, brk, x, l);
    }
    to_layout<CA, CB, BR>((0 - r) & 3, x, l);
    // the member addresses are formed again rather than kept live across
    // the phases
    const uint32_t *rows = a.rows;
    asm volatile("" : "+s"(rows));
    uint32_t col[NB];
    wfor<0, NB>([&](auto J) {
        constexpr int j = decltype(J)::value;
        col[j] = ld(rows, t.sB + j) * (kLanes * 8u) + lane8;
    });
    if (out)
        wfor<0, NA>([&](auto I) {
            constexpr int i = decltype(I)::value;
            double *ro = out + base + (long long)ld(rows, t.sA + i) * kLanes;
            wfor<0, NB>([&](auto J) { st_at(x[i * NB + decltype(J)::value], ro, col[decltype(J)::value]); });
        });
    if constexpr (BR)
        wfor<0, NA>([&](auto I) {
            constexpr int i = decltype(I)::value;
            double *rl = lam + base + (long long)ld(rows, t.sA + i) * kLanes;
            wfor<0, NB>([&](auto J) { st_at(l[i * NB + decltype(J)::value], rl, col[decltype(J)::value]); });
        });
}

// ------------------------------------------ vector-major (k_winv)
// Every phase of the pass on a tile held in registers, lane = tile (its
// outer bits in outer), angles uniform; the adjoint adds each operator's
// bracket to this lane's row in brk.
template <int CA, int CB, bool BR, bool TU>
__device__ __forceinline__ void phases_v(const WinArgs &a, uint32_t outer, double (&x)[C4[CA] * C4[CB]],
                                         double (&l)[BR ? C4[CA] * C4[CB] : 1], const double2 *__restrict__ cs,
                                         double *brk)
{
    const int lane = threadIdx.x & (kLanes - 1);
    const uint32_t *B = a.blob;
    int r = 0;
    for (int pi = 0; pi < a.nph; pi++) {
        const int p = BR ? a.nph - 1 - pi : pi;
        const int k0 = (int)ld(B, p * 8 + 0), nk = (int)ld(B, p * 8 + 1), kp = (int)ld(B, p * 8 + 3);
        const uint32_t spectM = ld(B, p * 8 + 6);
        const int lp = __builtin_ctz(ld(B, p * 8 + 4)) - a.w0;
        to_layout<CA, CB, BR>((lp - r) & 3, x, l);
        r = lp;
        auto one = [&](int kk) {
            const uint32_t m = ld(B, a.oMask + kp + kk) & spectM;
            wr_op_args o;
            o.cs = cs[k0 + kk];
            o.negT = __popc(m & outer) & 1;  // per lane
            o.mA = rot4((m >> (a.n + a.w0)) & 0xfu, lp);
            o.mB = rot4((m >> a.w0) & 0xfu, lp);
            o.progT = a.oProgT + (p * kOrbMaxOps + kk) * 16;
            return o;
        };
        if constexpr (TU) {
            wfor<0, 5>([&](auto J) {
                constexpr int kk = BR ? 4 - decltype(J)::value : decltype(J)::value;
                const double2 cst = cs[k0 + kk];
                const uint32_t sg = ld(B, a.oProgT + (p * kOrbMaxOps + kk) * 16 + 9);
                double br = 0;
                op_tups<CA, CB, kTupsKind[kk], BR>(cst.x, flip(cst.y, sg), x, l, br);
                if constexpr (BR) brk[(kp + kk) * kLanes + lane] += flip(br, sg);
            });
        } else {
            for (int jj = 0; jj < nk; jj++) {
                const int kk = BR ? nk - 1 - jj : jj;
                const wr_op_args o = one(kk);
                double br = 0;
                op_pair<CA, CB, 0, BR>(a, o.progT, o.negT, o.mA, o.mB, o.cs.x, o.cs.y, x, l, br);
                if constexpr (BR) brk[(kp + kk) * kLanes + lane] += br;
            }
        }
    }
    to_layout<CA, CB, BR>((0 - r) & 3, x, l);
}

// One vector per grid row (vector-major layout, vstride doubles apart);
// a wave takes a task of up to 64 tiles of one class (inner counts), lane
// = tile: the member indices are compile-time as above, the addresses and
// the outer-bit signs per lane, angles and program words uniform.  rec: the
// task's record block, word q of lane l at rec[q * 64 + l]: NA alpha rows,
// NB beta rows, the alpha and the beta outer bits.
template <int CA, int CB, bool BR, bool TU>
__device__ __forceinline__ void tile_v(const WinArgs &a, const uint32_t *__restrict__ rec, bool valid, double *v,
                                       double *lam, const double2 *__restrict__ cs, double *brk)
{
    constexpr int NA = C4[CA], NB = C4[CB], NM = NA * NB;
    const int lane = threadIdx.x & (kLanes - 1);
    uint32_t ra[NA], rb[NB];
    double x[NM], l[BR ? NM : 1];
    wfor<0, NA>([&](auto I) { ra[decltype(I)::value] = rec[decltype(I)::value * kLanes + lane]; });
    wfor<0, NB>([&](auto I) { rb[decltype(I)::value] = rec[(NA + decltype(I)::value) * kLanes + lane]; });
    const uint32_t outer = TU ? 0u : (rec[(NA + NB) * kLanes + lane] << a.n) | rec[(NA + NB + 1) * kLanes + lane];
    wfor<0, NM>([&](auto I) {
        constexpr int m = decltype(I)::value, i = m / NB, j = m % NB;
        const long long g = (long long)ra[i] + rb[j];
        x[m] = valid ? v[g] : 0.0;
        if constexpr (BR) l[m] = valid ? lam[g] : 0.0;
    });
    phases_v<CA, CB, BR, TU>(a, outer, x, l, cs, brk);
    wfor<0, NM>([&](auto I) {
        constexpr int m = decltype(I)::value, i = m / NB, j = m % NB;
        const long long g = (long long)ra[i] + rb[j];
        if (valid) {
            v[g] = x[m];
            if constexpr (BR) lam[g] = l[m];
        }
    });
}

}  // namespace wr

template <bool BR, bool TU>
__global__ void __launch_bounds__(4 * kLanes, BR ? 2 : 4)
    k_winv(WinArgs a, const int4 *__restrict__ tasks, int ntasks, const uint32_t *__restrict__ rec, double *v,
           double *lam, long long vstride, const double2 *__restrict__ cs, int M, double *__restrict__ partial,
           int slots)
{
    constexpr int NW = 4;
    extern __shared__ __align__(16) double lds[];
    const int lane = threadIdx.x & (kLanes - 1);
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / kLanes);
    const int vec = blockIdx.y;
    double *brk = lds + (size_t)wv * a.nkp * kLanes;  // adjoint: this wave's rows, one slot per lane
    if (BR)
        for (int e = 0; e < a.nkp; e++) brk[e * kLanes + lane] = 0.0;
    v += (long long)vec * vstride;
    if (BR) lam += (long long)vec * vstride;
    cs += (long long)vec * M;
    for (int it = blockIdx.x * NW + wv; it < ntasks; it += gridDim.x * NW) {
        const int4 tk = tasks[it];  // {class, record block offset, count}
        const int cls = __builtin_amdgcn_readfirstlane(tk.x), off = __builtin_amdgcn_readfirstlane(tk.y),
                  cnt = __builtin_amdgcn_readfirstlane(tk.z);
        const bool valid = lane < cnt;
        switch (cls) {
#define AE_WV_CASE(CA, CB) \
    case CA * 5 + CB: wr::tile_v<CA, CB, BR, TU>(a, rec + off, valid, v, lam, cs, brk); break;
            AE_WV_CASE(0, 0) AE_WV_CASE(0, 1) AE_WV_CASE(0, 2) AE_WV_CASE(0, 3) AE_WV_CASE(0, 4)
            AE_WV_CASE(1, 0) AE_WV_CASE(1, 1) AE_WV_CASE(1, 2) AE_WV_CASE(1, 3) AE_WV_CASE(1, 4)
            AE_WV_CASE(2, 0) AE_WV_CASE(2, 1) AE_WV_CASE(2, 2) AE_WV_CASE(2, 3) AE_WV_CASE(2, 4)
            AE_WV_CASE(3, 0) AE_WV_CASE(3, 1) AE_WV_CASE(3, 2) AE_WV_CASE(3, 3) AE_WV_CASE(3, 4)
            AE_WV_CASE(4, 0) AE_WV_CASE(4, 1) AE_WV_CASE(4, 2) AE_WV_CASE(4, 3) AE_WV_CASE(4, 4)
#undef AE_WV_CASE
        default: break;
        }
    }
    if (BR) {
        __syncthreads();
        // fixed-order sum over the waves, then over the lanes (tiles)
        for (int kp = wv; kp < a.nkp; kp += NW) {
            double s = 0;
#pragma unroll
            for (int q = 0; q < NW; q++) s += lds[((size_t)q * a.nkp + kp) * kLanes + lane];
            s = wave_sum(s);
            if (lane == 0) partial[((long long)vec * M + a.kpOp[kp]) * slots + blockIdx.x] = s;
        }
    }
}

// ------------------------------- vector-major, staged alpha rows (k_winvs)
namespace wr {
// A tile whose members are read from the block's staged alpha rows: member
// (i, j) at xs[i * nb + beta row j]; lane = tile of the wave's task, whose
// record block holds per lane NB beta rows (beta ranks) and the beta outer
// bits.
template <int CA, int CB, bool BR, bool TU>
__device__ __forceinline__ void tile_s(const WinArgs &a, const uint32_t *__restrict__ rec, bool valid,
                                       uint32_t outerA, double *xs, double *ls, int nb,
                                       const double2 *__restrict__ cs, double *brk)
{
    constexpr int NA = C4[CA], NB = C4[CB], NM = NA * NB;
    const int lane = threadIdx.x & (kLanes - 1);
    uint32_t rb[NB];
    double x[NM], l[BR ? NM : 1];
    wfor<0, NB>([&](auto J) { rb[decltype(J)::value] = rec[decltype(J)::value * kLanes + lane]; });
    const uint32_t outer = TU ? 0u : (outerA << a.n) | rec[NB * kLanes + lane];
    wfor<0, NM>([&](auto I) {
        constexpr int m = decltype(I)::value, i = m / NB, j = m % NB;
        const int idx = i * nb + (int)rb[j];
        x[m] = valid ? xs[idx] : 0.0;
        if constexpr (BR) l[m] = valid ? ls[idx] : 0.0;
    });
    phases_v<CA, CB, BR, TU>(a, outer, x, l, cs, brk);
    wfor<0, NM>([&](auto I) {
        constexpr int m = decltype(I)::value, i = m / NB, j = m % NB;
        const int idx = i * nb + (int)rb[j];
        if (valid) {
            xs[idx] = x[m];
            if constexpr (BR) ls[idx] = l[m];
        }
    });
}
}  // namespace wr

template <bool BR, bool TU>
__global__ void __launch_bounds__((BR ? 8 : 16) * kLanes)
    k_winvs(WinArgs a, const int4 *__restrict__ tasks, int ntasks, const uint32_t *__restrict__ rec, int nb,
            double *v, double *lam, long long vstride, const double2 *__restrict__ cs, int M,
            double *__restrict__ partial, int slots)
{
    extern __shared__ __align__(16) double lds[];
    const int lane = threadIdx.x & (kLanes - 1);
    const int nw = blockDim.x / kLanes;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / kLanes);
    const int vec = blockIdx.y;
    double *xs = lds, *ls = xs + (size_t)kWinVsRows * nb;
    double *brk = (BR ? ls + (size_t)kWinVsRows * nb : ls) + (size_t)wv * a.nkp * kLanes;
    if (BR)
        for (int e = 0; e < a.nkp; e++) brk[e * kLanes + lane] = 0.0;
    v += (long long)vec * vstride;
    if (BR) lam += (long long)vec * vstride;
    cs += (long long)vec * M;
    const uint32_t *Bl = a.blob;
    for (int ja = blockIdx.x; ja < a.nAcls; ja += gridDim.x) {
        const uint32_t fa = wr::ld(Bl, a.oCls + ja);
        const int ca = (int)(fa >> 24), na = ca == 0 || ca == 4 ? 1 : ca == 2 ? 6 : 4;
        const int sA = (int)wr::ld(Bl, a.oCStart + ja);
        __syncthreads();  // the previous class's rows are written back
        for (int i = 0; i < na; i++) {
            const long long row = wr::ld(a.rows, sA + i);
            for (int b = threadIdx.x; b < nb; b += blockDim.x) {
                xs[i * nb + b] = v[row + b];
                if (BR) ls[i * nb + b] = lam[row + b];
            }
        }
        __syncthreads();
        for (int it = wv; it < ntasks; it += nw) {
            const int4 tk = tasks[it];  // {beta inner count, record block offset, count}
            const int cb = __builtin_amdgcn_readfirstlane(tk.x), off = __builtin_amdgcn_readfirstlane(tk.y),
                      cnt = __builtin_amdgcn_readfirstlane(tk.z);
            const bool valid = lane < cnt;
            switch (ca * 5 + cb) {
#define AE_WS_CASE(CA, CB)                                                                                          \
    case CA * 5 + CB:                                                                                               \
        wr::tile_s<CA, CB, BR, TU>(a, rec + off, valid, fa & 0xffffffu, xs, ls, nb, cs, brk);                       \
        break;
                AE_WS_CASE(0, 0) AE_WS_CASE(0, 1) AE_WS_CASE(0, 2) AE_WS_CASE(0, 3) AE_WS_CASE(0, 4)
                AE_WS_CASE(1, 0) AE_WS_CASE(1, 1) AE_WS_CASE(1, 2) AE_WS_CASE(1, 3) AE_WS_CASE(1, 4)
                AE_WS_CASE(2, 0) AE_WS_CASE(2, 1) AE_WS_CASE(2, 2) AE_WS_CASE(2, 3) AE_WS_CASE(2, 4)
                AE_WS_CASE(3, 0) AE_WS_CASE(3, 1) AE_WS_CASE(3, 2) AE_WS_CASE(3, 3) AE_WS_CASE(3, 4)
                AE_WS_CASE(4, 0) AE_WS_CASE(4, 1) AE_WS_CASE(4, 2) AE_WS_CASE(4, 3) AE_WS_CASE(4, 4)
#undef AE_WS_CASE
            default: break;
            }
        }
        __syncthreads();
        for (int i = 0; i < na; i++) {
            const long long row = wr::ld(a.rows, sA + i);
            for (int b = threadIdx.x; b < nb; b += blockDim.x) {
                v[row + b] = xs[i * nb + b];
                if (BR) lam[row + b] = ls[i * nb + b];
            }
        }
    }
    if (BR) {
        __syncthreads();
        double *rows0 = ls + (size_t)kWinVsRows * nb;
        for (int kp = wv; kp < a.nkp; kp += nw) {
            double s = 0;
            for (int q = 0; q < nw; q++) s += rows0[((size_t)q * a.nkp + kp) * kLanes + lane];
            s = wave_sum(s);
            if (lane == 0) partial[((long long)vec * M + a.kpOp[kp]) * slots + blockIdx.x] = s;
        }
    }
}

#define AE_WINVS(BR, TU)                                                                                       \
    template __global__ void k_winvs<BR, TU>(WinArgs, const int4 *__restrict__, int, const uint32_t *__restrict__, \
                                             int, double *, double *, long long, const double2 *__restrict__, int, \
                                             double *__restrict__, int);
AE_WINVS(false, false)
AE_WINVS(true, false)
AE_WINVS(false, true)
AE_WINVS(true, true)
#undef AE_WINVS

#define AE_WINV(BR, TU)                                                                                      \
    template __global__ void k_winv<BR, TU>(WinArgs, const int4 *__restrict__, int, const uint32_t *__restrict__, \
                                            double *, double *, long long, const double2 *__restrict__, int,      \
                                            double *__restrict__, int);
AE_WINV(false, false)
AE_WINV(true, false)
AE_WINV(false, true)
AE_WINV(true, true)
#undef AE_WINV

// occupancy: forward 4 waves per SIMD (at most 128 VGPRs), adjoint 2 (256)
template <bool BR, bool BC, int NW, bool TU>

__global__ void __launch_bounds__(NW * kLanes, BR ? 2 : 4)
    k_winr(WinArgs a, const double *in, double *out, double *lam, long long cstride, const double2 *__restrict__ cs,
           int M, double *__restrict__ partial, int slots, double *__restrict__ epart)
{
    extern __shared__ __align__(16) double lds[];
    const int lane = threadIdx.x & (kLanes - 1);
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / kLanes);
    const int c = blockIdx.y;
    double *brk = lds + (size_t)wv * a.nkp * kLanes;  // adjoint: this wave's bracket rows
    double2 *csL = (double2 *)lds;                      // forward: the pass's angles
    if (BR) {
        for (int e = 0; e < a.nkp; e++) brk[e * kLanes + lane] = 0.0;
    } else {
        for (int e = wv; e < a.nkp; e += NW) csL[e * kLanes + lane] = cs[((long long)c * M + a.kpOp[e]) * kLanes + lane];
        __syncthreads();
    }
    const long long base = (long long)c * cstride;  // the chunk (lanes: the members' column offsets)
    const uint32_t *B = a.blob;
    const bool en = BR && epart != nullptr;
    double e = 0;  // en: this lane's (vector's) <x|l> over the wave's tiles
    for (int it = blockIdx.x * NW + wv; it < a.ntiles; it += gridDim.x * NW) {
        const int tl = __builtin_amdgcn_readfirstlane(a.tiles[it]);
        const int ja = tl / a.nBcls, jb = tl - ja * a.nBcls;
        const uint32_t fa = wr::ld(B, a.oCls + ja), fb = wr::ld(B, a.oCls + a.nAcls + jb);
        wr::TileCtx t;
        t.aOut = fa & 0xffffffu;
        t.bOut = fb & 0xffffffu;
        t.sA = (int)wr::ld(B, a.oCStart + ja);
        t.sB = a.nrowsA + (int)wr::ld(B, a.oCStart + a.nAcls + jb);
        const int cls = (int)(fa >> 24) * 5 + (int)(fb >> 24);
        switch (cls) {
#define AE_WR_CASE(CA, CB) \
    case CA * 5 + CB: wr::tile<CA, CB, BR, BC, NW, TU>(a, t, in, out, lam, base, cs, csL, c, M, brk, en, e); break;
            AE_WR_CASE(0, 0) AE_WR_CASE(0, 1) AE_WR_CASE(0, 2) AE_WR_CASE(0, 3) AE_WR_CASE(0, 4)
            AE_WR_CASE(1, 0) AE_WR_CASE(1, 1) AE_WR_CASE(1, 2) AE_WR_CASE(1, 3) AE_WR_CASE(1, 4)
            AE_WR_CASE(2, 0) AE_WR_CASE(2, 1) AE_WR_CASE(2, 2) AE_WR_CASE(2, 3) AE_WR_CASE(2, 4)
            AE_WR_CASE(3, 0) AE_WR_CASE(3, 1) AE_WR_CASE(3, 2) AE_WR_CASE(3, 3) AE_WR_CASE(3, 4)
            AE_WR_CASE(4, 0) AE_WR_CASE(4, 1) AE_WR_CASE(4, 2) AE_WR_CASE(4, 3) AE_WR_CASE(4, 4)
#undef AE_WR_CASE
        default: break;
        }
    }
    if (BR) {
        __syncthreads();
        for (int kp = wv; kp < a.nkp; kp += NW) {
            double s = 0;
#pragma unroll
            for (int q = 0; q < NW; q++) s += lds[((size_t)q * a.nkp + kp) * kLanes + lane];
            partial[(((long long)c * M + a.kpOp[kp]) * slots + blockIdx.x) * kLanes + lane] = s;
        }
        if (en) {
            // one energy row per block: P = 2^min(4, trailing zeros of the
            // grid) row groups of gridDim.x / P rows, block b in group b % P
            __shared__ double esum[NW][kLanes];
            esum[wv][lane] = e;
            __syncthreads();
            if (wv == 0) {
                double s = 0;
#pragma unroll
                for (int q = 0; q < NW; q++) s += esum[q][lane];
                const int P = 1 << min(4, __builtin_ctz(gridDim.x)), R = gridDim.x / P;
                const int k = blockIdx.x % P, r = blockIdx.x / P;
                epart[(((long long)c * P + k) * R + r) * kLanes + lane] = s;
            }
        }
    }
}

// ------------------------------------------- Hessian tangents (k_wint)
namespace wr {
// One TUPS operator on the tangent tile: psi, lambda and the lane's tangent
// rotate (S carries the operator's phase, scs = its scale times the phase);
// the injection lane adds scs G psi; xacc gets <lambda|G|b> without the
// phase, dacc <lambda|G G|psi>.
template <int CA, int CB, int KIND, int NM>
__device__ __forceinline__ void op_tan(double C, double S, double scs, bool inj, double (&b)[NM], double (&ps)[NM],
                                       double (&lm)[NM], double &xacc, double &dacc)
{
    constexpr int NB = C4[CB];
    constexpr int NSA = nsub(CA, 0), NSB = nsub(CB, 0);
    constexpr int PP[6] = {0, 0, 0, 1, 1, 2}, QQ[6] = {1, 2, 3, 2, 3, 3};
    wfor<0, NSA * NSB>([&](auto I) {
        constexpr Sub A = sub(CA, 0, decltype(I)::value / NSB);
        constexpr Sub Bs = sub(CB, 0, decltype(I)::value % NSB);
        constexpr unsigned slots = tups_slots(KIND, A.sh, Bs.sh);
        wfor<0, 6>([&](auto R) {
            constexpr int r = decltype(R)::value;
            if constexpr ((slots >> r) & 1u) {
                constexpr int p = PP[r], q = QQ[r];
                constexpr int mp = ((p >> 1) ? A.i1 : A.i0) * NB + ((p & 1) ? Bs.i1 : Bs.i0);
                constexpr int mq = ((q >> 1) ? A.i1 : A.i0) * NB + ((q & 1) ? Bs.i1 : Bs.i0);
                rot(ps[mp], ps[mq], C, S);
                rot(lm[mp], lm[mq], C, S);
                rot(b[mp], b[mq], C, S);
                if (inj) {
                    b[mp] = fma(scs, ps[mq], b[mp]);
                    b[mq] = fma(-scs, ps[mp], b[mq]);
                }
                xacc = fma(lm[mp], b[mq], xacc);
                xacc = fma(-lm[mq], b[mp], xacc);
                dacc = fma(-lm[mp], ps[mp], dacc);
                dacc = fma(-lm[mq], ps[mq], dacc);
            }
        });
    });
}

// lambda alone (a tile where psi and the tangents are zero)
template <int CA, int CB, int KIND, int NM>
__device__ __forceinline__ void op_lam(double C, double S, double (&lm)[NM])
{
    constexpr int NB = C4[CB];
    constexpr int NSA = nsub(CA, 0), NSB = nsub(CB, 0);
    constexpr int PP[6] = {0, 0, 0, 1, 1, 2}, QQ[6] = {1, 2, 3, 2, 3, 3};
    wfor<0, NSA * NSB>([&](auto I) {
        constexpr Sub A = sub(CA, 0, decltype(I)::value / NSB);
        constexpr Sub Bs = sub(CB, 0, decltype(I)::value % NSB);
        constexpr unsigned slots = tups_slots(KIND, A.sh, Bs.sh);
        wfor<0, 6>([&](auto R) {
            constexpr int r = decltype(R)::value;
            if constexpr ((slots >> r) & 1u) {
                constexpr int p = PP[r], q = QQ[r];
                constexpr int mp = ((p >> 1) ? A.i1 : A.i0) * NB + ((p & 1) ? Bs.i1 : Bs.i0);
                constexpr int mq = ((q >> 1) ? A.i1 : A.i0) * NB + ((q & 1) ? Bs.i1 : Bs.i0);
                rot(lm[mp], lm[mq], C, S);
            }
        });
    });
}

// A tile of the pass that psi cannot have reached: psi and the tangents are
// zero on it before and after, so only lambda is carried (chunk 0, written
// by lane 0) and psi's zeros are written to psiOut.
template <int CA, int CB>
__device__ __forceinline__ void tile_l(const WinArgs &a, const TileCtx &t, const double *lamIn, double *psiOut,
                                       double *lamOut, const double2 *__restrict__ cs)
{
    constexpr int NA = C4[CA], NB = C4[CB], NM = NA * NB;
    const int lane = threadIdx.x & (kLanes - 1);
    const uint32_t *Bl = a.blob;
    double lm[NM], none[1];
    uint32_t colS[NB];
    wfor<0, NB>([&](auto J) { colS[decltype(J)::value] = ld(a.rows, t.sB + decltype(J)::value) * 8u; });
    wfor<0, NA>([&](auto I) {
        constexpr int i = decltype(I)::value;
        const double *rl = lamIn + (long long)ld(a.rows, t.sA + i);
        wfor<0, NB>([&](auto J) {
            lm[i * NB + decltype(J)::value] = *(const double *)((const char *)rl + colS[decltype(J)::value]);
        });
    });
    int r = 0;
    for (int p = 0; p < a.nph; p++) {
        const int lp = __builtin_ctz(ld(Bl, p * 8 + 4)) - a.w0;
        to_layout<CA, CB, false>((lp - r) & 3, lm, none);
        r = lp;
        const int k0 = (int)ld(Bl, p * 8 + 0);
        wfor<0, 5>([&](auto J) {
            constexpr int kk = decltype(J)::value;
            const double2 cst = cs[k0 + kk];
            const uint32_t sg = ld(Bl, a.oProgT + (p * kOrbMaxOps + kk) * 16 + 9);
            op_lam<CA, CB, kTupsKind[kk]>(cst.x, flip(cst.y, sg), lm);
        });
    }
    to_layout<CA, CB, false>((0 - r) & 3, lm, none);
    if (lane == 0)
        wfor<0, NA>([&](auto I) {
            constexpr int i = decltype(I)::value;
            const long long ra = ld(a.rows, t.sA + i);
            double *rp = psiOut + ra, *rl = lamOut + ra;
            wfor<0, NB>([&](auto J) {
                constexpr int j = decltype(J)::value;
                *(double *)((char *)rp + colS[j]) = 0.0;
                *(double *)((char *)rl + colS[j]) = lm[i * NB + j];
            });
        });
}

// A tile of the tangent pass: lane = tangent (chunk c), psi and lambda one
// vector each (every lane holds the same values).  bx: this wave's bracket
// rows (nkp x 64), bd: its nkp D sums.
template <int CA, int CB>
__device__ __forceinline__ void tile_t(const WinArgs &a, const TileCtx &t, const double *psiIn, const double *lamIn,
                                       double *psiOut, double *lamOut, double *B, const double2 *__restrict__ cs,
                                       const int *__restrict__ inj, const double *__restrict__ scale, int c, double *bx,
                                       double *bd)
{
    constexpr int NA = C4[CA], NB = C4[CB], NM = NA * NB;
    const int lane = threadIdx.x & (kLanes - 1);
    const uint32_t lane8 = lane * 8u;
    const uint32_t *Bl = a.blob;
    double b[NM], ps[NM], lm[NM];
    {
        uint32_t col[NB], colS[NB];
        wfor<0, NB>([&](auto J) {
            constexpr int j = decltype(J)::value;
            const uint32_t rb = ld(a.rows, t.sB + j);
            col[j] = rb * (kLanes * 8u) + lane8;
            colS[j] = rb * 8u;
        });
        wfor<0, NA>([&](auto I) {
            constexpr int i = decltype(I)::value;
            const long long ra = ld(a.rows, t.sA + i);
            const double *rb_ = B + ra * kLanes, *rp = psiIn + ra, *rl = lamIn + ra;
            wfor<0, NB>([&](auto J) {
                constexpr int j = decltype(J)::value;
                b[i * NB + j] = ld_at(rb_, col[j]);
                ps[i * NB + j] = *(const double *)((const char *)rp + colS[j]);
                lm[i * NB + j] = *(const double *)((const char *)rl + colS[j]);
            });
        });
    }
    double none[1];
    int r = 0;
    for (int p = 0; p < a.nph; p++) {
        const int lp = __builtin_ctz(ld(Bl, p * 8 + 4)) - a.w0;
        to_layout<CA, CB, true>((lp - r) & 3, b, ps);
        to_layout<CA, CB, false>((lp - r) & 3, lm, none);
        r = lp;
        const int k0 = (int)ld(Bl, p * 8 + 0), kp = (int)ld(Bl, p * 8 + 3);
        wfor<0, 5>([&](auto J) {
            constexpr int kk = decltype(J)::value;
            const int k = k0 + kk;
            const double2 cst = cs[k];
            const uint32_t sg = ld(Bl, a.oProgT + (p * kOrbMaxOps + kk) * 16 + 9);
            const bool in = __builtin_amdgcn_readfirstlane(inj[k]) == c * kLanes + lane;
            double xacc = 0, dacc = 0;
            op_tan<CA, CB, kTupsKind[kk]>(cst.x, flip(cst.y, sg), flip(scale[k], sg), in, b, ps, lm, xacc, dacc);
            bx[(kp + kk) * kLanes + lane] += flip(xacc, sg);
            if (lane == 0) bd[kp + kk] += dacc;
        });
    }
    to_layout<CA, CB, true>((0 - r) & 3, b, ps);
    to_layout<CA, CB, false>((0 - r) & 3, lm, none);
    const uint32_t *rows = a.rows;
    asm volatile("" : "+s"(rows));
    uint32_t col[NB], colS[NB];
    wfor<0, NB>([&](auto J) {
        constexpr int j = decltype(J)::value;
        const uint32_t rb = ld(rows, t.sB + j);
        col[j] = rb * (kLanes * 8u) + lane8;
        colS[j] = rb * 8u;
    });
    wfor<0, NA>([&](auto I) {
        constexpr int i = decltype(I)::value;
        const long long ra = ld(rows, t.sA + i);
        double *rb_ = B + ra * kLanes;
        wfor<0, NB>([&](auto J) { st_at(b[i * NB + decltype(J)::value], rb_, col[decltype(J)::value]); });
        if (c == 0 && lane == 0) {
            double *rp = psiOut + ra, *rl = lamOut + ra;
            wfor<0, NB>([&](auto J) {
                constexpr int j = decltype(J)::value;
                *(double *)((char *)rp + colS[j]) = ps[i * NB + j];
                *(double *)((char *)rl + colS[j]) = lm[i * NB + j];
            });
        }
    });
}
}  // namespace wr

__global__ void __launch_bounds__(2 * kLanes, 2)
    k_wint(WinArgs a, const double *psiIn, const double *lamIn, double *psiOut, double *lamOut, double *B,
           long long cstride, const double2 *__restrict__ cs, int M, const int *__restrict__ inj,
           const double *__restrict__ scale, double *__restrict__ xpart, double *__restrict__ dpart, int slots,
           int nreach)
{
    constexpr int NW = 2;
    extern __shared__ __align__(16) double lds[];
    const int lane = threadIdx.x & (kLanes - 1);
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / kLanes);
    const int c = blockIdx.y;
    const int rowsW = a.nkp * (kLanes + 1);  // doubles per wave: X rows, then D
    double *bx = lds + (size_t)wv * rowsW, *bd = bx + (size_t)a.nkp * kLanes;
    for (int e = lane; e < rowsW; e += kLanes) bx[e] = 0.0;
    B += (long long)c * cstride;
    const uint32_t *Bl = a.blob;
    // the first nreach listed tiles are the ones psi can have reached; the
    // rest carry lambda only (chunk 0)
    const int nt = c == 0 ? a.ntiles : nreach;
    for (int it = blockIdx.x * NW + wv; it < nt; it += gridDim.x * NW) {
        const int tl = __builtin_amdgcn_readfirstlane(a.tiles[it]);
        const int ja = tl / a.nBcls, jb = tl - ja * a.nBcls;
        const uint32_t fa = wr::ld(Bl, a.oCls + ja), fb = wr::ld(Bl, a.oCls + a.nAcls + jb);
        wr::TileCtx t;
        t.aOut = fa & 0xffffffu;
        t.bOut = fb & 0xffffffu;
        t.sA = (int)wr::ld(Bl, a.oCStart + ja);
        t.sB = a.nrowsA + (int)wr::ld(Bl, a.oCStart + a.nAcls + jb);
        const int cls = (int)(fa >> 24) * 5 + (int)(fb >> 24) + (it < nreach ? 0 : 25);
        switch (cls) {
#define AE_WT_CASE(CA, CB) \
    case CA * 5 + CB: wr::tile_t<CA, CB>(a, t, psiIn, lamIn, psiOut, lamOut, B, cs, inj, scale, c, bx, bd); break; \
    case 25 + CA * 5 + CB: wr::tile_l<CA, CB>(a, t, lamIn, psiOut, lamOut, cs); break;
            AE_WT_CASE(0, 0) AE_WT_CASE(0, 1) AE_WT_CASE(0, 2) AE_WT_CASE(0, 3) AE_WT_CASE(0, 4)
            AE_WT_CASE(1, 0) AE_WT_CASE(1, 1) AE_WT_CASE(1, 2) AE_WT_CASE(1, 3) AE_WT_CASE(1, 4)
            AE_WT_CASE(2, 0) AE_WT_CASE(2, 1) AE_WT_CASE(2, 2) AE_WT_CASE(2, 3) AE_WT_CASE(2, 4)
            AE_WT_CASE(3, 0) AE_WT_CASE(3, 1) AE_WT_CASE(3, 2) AE_WT_CASE(3, 3) AE_WT_CASE(3, 4)
            AE_WT_CASE(4, 0) AE_WT_CASE(4, 1) AE_WT_CASE(4, 2) AE_WT_CASE(4, 3) AE_WT_CASE(4, 4)
#undef AE_WT_CASE
        default: break;
        }
    }
    __syncthreads();
    for (int kp = wv; kp < a.nkp; kp += NW) {
        double sx = 0, sd = 0;
#pragma unroll
        for (int q = 0; q < NW; q++) {
            sx += lds[(size_t)q * rowsW + (size_t)kp * kLanes + lane];
            sd += lds[(size_t)q * rowsW + (size_t)a.nkp * kLanes + kp];
        }
        const int k = a.kpOp[kp];
        xpart[(((long long)c * M + k) * slots + blockIdx.x) * kLanes + lane] = sx;
        if (c == 0 && lane == 0) dpart[(long long)k * slots + blockIdx.x] = sd;
    }
}

#define AE_WINR(BR, BC, NW, TU)                                                                             \
    template __global__ void k_winr<BR, BC, NW, TU>(WinArgs, const double *, double *, double *, long long,\
                                                    const double2 *__restrict__, int, double *__restrict__, int,   \
                                                    double *__restrict__);
AE_WINR(false, false, 4, false)
AE_WINR(false, true, 4, false)
AE_WINR(true, false, 2, false)
AE_WINR(false, false, 4, true)
AE_WINR(false, true, 4, true)
AE_WINR(true, false, 2, true)
#undef AE_WINR

} // namespace ae
