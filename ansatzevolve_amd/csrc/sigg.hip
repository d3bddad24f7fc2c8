// Lane-layout sparse sigma at 12 orbitals with grouped targets (gfx950).
//
// Same terms, block mapping, staging and coefficient row as k_sigw_sparse
// (the alpha-beta doubles and the beta-diagonal part of every alpha single's
// source row, the alpha-dependent spectator part of the beta singles through
// the identity row, and the diagonal; src/hamiltonianmatrix.cpp:541-640), but
// the targets of a 16-lane row are all beta strings (bL, h) with one upper
// half h (orbitals 6..11), so a source read from LDS feeds every target of
// the row that it reaches.  The row programs, per-row source bases and
// coefficient-slot tables are generated (sigg_prog.inc, from sigg.py): a
// read is ds_read_b64 at (row base k) + static offset; an FMA takes its
// coefficient from lane c mod 16 of chunk c / 16 of the row (row_newbcast),
// with a static sign as the source negate modifier.
#include "kernels.hpp"

namespace ae {
namespace {

typedef __attribute__((address_space(3))) double sg_lds_f64;
__device__ __forceinline__ double lds_at(uint32_t byteaddr) { return *(const sg_lds_f64 *)(size_t)byteaddr; }
__device__ __forceinline__ double sg_thr13(double v) { return fabs(v) > 1e-13 ? v : 0.0; }
__device__ __forceinline__ int sg_between_parity(uint32_t s, int r, int q)
{
    const int lo = r < q ? r : q, hi = r < q ? q : r;
    const uint32_t m = ((1u << hi) - 1) & ~((1u << (lo + 1)) - 1);
    return __popc(s & m) & 1;
}

// acc += (c of lane J of this 16-lane row) * x, and the same with -c; c is
// not written by a VALU instruction within two instructions of its use (the
// coefficient chunks pass through an s_nop after they are formed)
template <int J>
__device__ __forceinline__ void fma_row_bcast(double &acc, double c, double x)
{
    asm("v_fmac_f64_dpp %0, %1, %2 row_newbcast:%3 row_mask:0xf bank_mask:0xf bound_ctrl:1"
        : "+v"(acc)
        : "v"(c), "v"(x), "i"(J));
}
template <int J>
__device__ __forceinline__ void fma_row_bcast_neg(double &acc, double c, double x)
{
    asm("v_fmac_f64_dpp %0, -%1, %2 row_newbcast:%3 row_mask:0xf bank_mask:0xf bound_ctrl:1"
        : "+v"(acc)
        : "v"(c), "v"(x), "i"(J));
}

constexpr int kSiggAcc = 20;  // targets of the largest group, C(6, 3)

#include "sigg_prog.inc"

template <int K>
struct SiggTables;
template <>
struct SiggTables<6> {
    static __device__ __forceinline__ const uint32_t *base() { return kSiggBase6; }
    static __device__ __forceinline__ const uint32_t *ent() { return kSiggEnt6; }
    static __device__ __forceinline__ const int *plan() { return kSiggPlan6; }
};

template <int M, int AO>
__device__ __forceinline__ void sigg_prog6(double (&acc)[kSiggAcc], const uint32_t (&ad)[16], const double (&cv)[6])
{
    if constexpr (M == 0) sigg_prog_6_0<AO>(acc, ad, cv);
    if constexpr (M == 1) sigg_prog_6_1<AO>(acc, ad, cv);
    if constexpr (M == 2) sigg_prog_6_2<AO>(acc, ad, cv);
    if constexpr (M == 3) sigg_prog_6_3<AO>(acc, ad, cv);
    if constexpr (M == 4) sigg_prog_6_4<AO>(acc, ad, cv);
    if constexpr (M == 5) sigg_prog_6_5<AO>(acc, ad, cv);
    if constexpr (M == 6) sigg_prog_6_6<AO>(acc, ad, cv);
}

constexpr int sg_comb6(int m) { return m == 0 || m == 6 ? 1 : m == 1 || m == 5 ? 6 : m == 2 || m == 4 ? 15 : 20; }

// one task of this wave for the staged source row: its row's bases and
// coefficient chunks, then its program
template <int K, int M, int AO>
__device__ __forceinline__ void sigg_task(int h, uint32_t lb, uint32_t gsb, const double *dt,
                                          const double *__restrict__ diag, bool ident, double (&acc)[kSiggAcc])
{
    int hr = h < 0 ? 0 : h;  // padding rows run row 0's tables; their sums are never stored
    // the tables are read per source row: hoisted out of the row loop they
    // would hold 22 registers for the whole kernel
    asm volatile("" : "+v"(hr));
    const int v = threadIdx.x & 15;
    const uint32_t *bt = SiggTables<K>::base() + hr * 16;
    const uint32_t *et = SiggTables<K>::ent() + hr * 96 + v;
    uint32_t ad[16];
#pragma unroll
    for (int k = 0; k < 16; k++) ad[k] = lb + bt[k];
    double cv[6];
#pragma unroll
    for (int c = 0; c < 6; c++) {
        const uint32_t e = et[16 * c];
        if (e & 0x80000000u) {
            // diagonal slot: occupancy sums of the target string, or for the
            // identity row the diagonal element
            const uint32_t s = e & 0xfffu, rk = (e >> 12) & 0x7ffffu;
            cv[c] = ident ? diag[rk] : dt[s & 63] + dt[64 + ((s >> 6) & 63)];
        } else {
            cv[c] = lds_at(gsb + e);
        }
    }
    asm volatile("s_nop 1" : "+v"(cv[0]), "+v"(cv[1]), "+v"(cv[2]), "+v"(cv[3]), "+v"(cv[4]), "+v"(cv[5]));
    if constexpr (K == 6) sigg_prog6<M, AO>(acc, ad, cv);
}

// targets b0 .. b0 + NT - 1 of alpha row ar; with z (exchange-symmetric
// batches) both same-spin parts are added: z[ar][b] and, transposed,
// z[b][ar]
template <int AO, int NT>
__device__ __forceinline__ void sigg_store(const double (&acc)[kSiggAcc], double *__restrict__ y,
                                           const double *__restrict__ z, int ar, int nb, int b0, int v)
{
    const long long row0 = (long long)ar * nb + b0;
    if (z) {
        double zs[NT];
#pragma unroll
        for (int t = 0; t < NT; t++)
            zs[t] = z[(row0 + t) * kLanes + v] + z[((long long)(b0 + t) * nb + ar) * kLanes + v];
#pragma unroll
        for (int t = 0; t < NT; t++) y[(row0 + t) * kLanes + v] = acc[AO + t] + zs[t];
    } else {
#pragma unroll
        for (int t = 0; t < NT; t++) y[(row0 + t) * kLanes + v] = acc[AO + t];
    }
}

// The whole source-row loop for a wave whose tasks are programs MA (and MB,
// accumulators from 10; -1 none).  Each kind of wave runs its own copy of the
// loop, so each keeps its accumulators in fixed registers; all copies pass
// the same barriers the same number of times.
template <int K, int MA, int MB>
__device__ __forceinline__ void sigg_loop(const SigvArgs &a, int ar, int hA, int hB, double *gs, double *dt,
                                          double *xs, const double *__restrict__ x, double *__restrict__ y,
                                          const double *__restrict__ z, bool runA, bool runB)
{
    const MfArgs &m = a.m;
    const int n = m.n, nb = m.nb, nn = n * n;
    const int tid = threadIdx.x, lane = tid & (kLanes - 1), v = lane & 15;
    const int wu = __builtin_amdgcn_readfirstlane(tid / kLanes);
    const uint32_t alpha = m.unA[ar];
    const int ka = __popc(alpha), ne = n - ka;
    const int R = 1 + ka * ne;
    const uint32_t gsb = (uint32_t)(size_t)(sg_lds_f64 *)gs;
    const uint32_t lb = (uint32_t)(size_t)(sg_lds_f64 *)xs + v * 8;
    const double *diag = m.diag + (long long)ar * nb;
    double acc[kSiggAcc];
#pragma unroll
    for (int i = 0; i < kSiggAcc; i++) acc[i] = 0;
    for (int r = 0; r < R; r++) {
        // source row r in ascending rank: the identity row (J = ar) or the
        // alpha single (occupied rA, empty qA) with its coefficients
        const int4 se = a.src[(long long)ar * R + r];
        const int J = se.x, rA = se.y < 0 ? 0 : se.y, qA = se.z < 0 ? 0 : se.z;
        const bool ident = se.y < 0;
        __syncthreads();  // every read of the previous row is done
        {
            // the source row into LDS by asynchronous global-to-LDS copies, all
            // issued before any wait: 8 betas (1 KB) per wave instruction
            const double *src = x + (long long)J * nb * kLanes + (lane & 7) * 2;
            for (int c = wu; c * 8 < nb; c += kSigWThreads / kLanes) {
                const int beta = c * 8 + (lane >> 3);
                __builtin_amdgcn_global_load_lds((const void *)(src + (long long)(beta < nb ? beta : 0) * kLanes),
                                                 (__attribute__((address_space(3))) void *)(xs + c * 128), 16, 0,
                                                 0);
            }
        }
        if (ident) {
            // identity row: J_alpha[e] = sum over occupied alpha o of g[oo][e] (r != q only)
            for (int e = tid; e < nn; e += kSigWThreads) {
                double vj = 0;
                if (e / n != e % n)
                    for (uint32_t o = alpha; o; o &= o - 1) {
                        const int xo = __ffs(o) - 1;
                        vj += sg_thr13(m.g[(size_t)(xo * n + xo) * nn + e]);
                    }
                gs[e] = vj;
                gs[kSigNN + e] = -vj;
            }
        } else {
            const bool neg = sg_between_parity(alpha, rA, qA);
            const double *gb = m.g + (size_t)(rA * n + qA) * nn;
            for (int e = tid; e < nn; e += kSigWThreads) {
                const double cf = neg ? -sg_thr13(gb[e]) : sg_thr13(gb[e]);
                gs[e] = cf;
                gs[kSigNN + e] = -cf;
            }
        }
        if (tid < kSigGS - 2 * kSigNN) gs[2 * kSigNN + tid] = 0.0;
        // beta-diagonal sums over the occupied orbitals of a string, by its
        // low and high 6 bits (zero for the identity row)
        if (tid < 128) {
            const int hi = tid >> 6, t = tid & 63;
            double d = 0;
            if (!ident) {
                const bool neg = sg_between_parity(alpha, rA, qA);
                const double *gb = m.g + (size_t)(rA * n + qA) * nn;
                double gv[6];
#pragma unroll
                for (int b = 0; b < 6; b++) gv[b] = gb[(b + 6 * hi) * (n + 1)];
#pragma unroll
                for (int b = 0; b < 6; b++)
                    if ((t >> b) & 1) d += neg ? -sg_thr13(gv[b]) : sg_thr13(gv[b]);
            }
            dt[tid] = d;
        }
        __builtin_amdgcn_s_waitcnt(0x0f70);  // vmcnt(0): this thread's row copies have landed
        __syncthreads();
        if (runA) sigg_task<K, MA, 0>(hA, lb, gsb, dt, diag, ident, acc);
        if constexpr (MB >= 0)
            if (runB) sigg_task<K, MB, 10>(hB, lb, gsb, dt, diag, ident, acc);
    }
    const uint32_t *bt = SiggTables<K>::base();
    if (runA && hA >= 0) sigg_store<0, sg_comb6(MA)>(acc, y, z, ar, nb, (int)(bt[hA * 16] / 128), v);
    if constexpr (MB >= 0)
        if (runB && hB >= 0) sigg_store<10, sg_comb6(MB)>(acc, y, z, ar, nb, (int)(bt[hB * 16] / 128), v);
}

}  // namespace

template <int K>
__global__ void __launch_bounds__(kSigWThreads) k_sigg_sparse(SigvArgs a)
{
    extern __shared__ double sm[];
    // Block order: every alpha row of sub-chunk 0 (vectors 0-15), then of
    // sub-chunk 1, ...: a sub-chunk's 128-byte lines of the chunk (N x 128 B,
    // 109 MB at 12 orbitals) stay in the 256 MB last-level cache while its 37
    // readers per source row stage them.  Rows are spread over the 8 XCDs.
    const int bid = blockIdx.x, xcd = bid & 7, k = bid >> 3, nr = (a.m.na + 7) / 8;
    const int sub = k / nr, ar = xcd + 8 * (k - sub * nr);
    if (ar >= a.m.na) return;
    const long long coff = (long long)blockIdx.y * a.cstride + sub * kSigWVC;
    const double *__restrict__ x = a.x + coff;
    double *__restrict__ y = a.y + coff;
    const double *__restrict__ z = a.z ? a.z + coff : nullptr;
    double *gs = sm, *dt = sm + kSigGS, *xs = dt + 128;
    const int rs = (threadIdx.x & (kLanes - 1)) >> 4;
    const int wu = __builtin_amdgcn_readfirstlane(threadIdx.x / kLanes);
    // this wave's tasks: (m, accumulator offset, four rows) twice
    const int *pl = SiggTables<K>::plan() + wu * 12;
    const int mA = pl[0], hA = pl[2 + rs], mB = pl[6], hB = pl[8 + rs];
    static_assert(K == 6, "programs are generated for 6 beta electrons");
    // exchange-symmetric batches need the targets b >= ar only: a task runs
    // when one of its rows reaches that far (wave-uniform)
    bool runA = true, runB = true;
    if (z) {
        const uint32_t *bt = SiggTables<K>::base();
        const bool topA = hA >= 0 && (int)(bt[hA * 16] / 128) + sg_comb6(mA) - 1 >= ar;
        const bool topB = mB >= 0 && hB >= 0 && (int)(bt[hB * 16] / 128) + sg_comb6(mB) - 1 >= ar;
        runA = __ballot(topA) != 0;
        runB = __ballot(topB) != 0;
    }
    if (mA == 3) sigg_loop<K, 3, -1>(a, ar, hA, hB, gs, dt, xs, x, y, z, runA, runB);
    else if (mA == 2) sigg_loop<K, 2, -1>(a, ar, hA, hB, gs, dt, xs, x, y, z, runA, runB);
    else if (mA == 4) sigg_loop<K, 4, -1>(a, ar, hA, hB, gs, dt, xs, x, y, z, runA, runB);
    else if (mA == 5 && mB == 5) sigg_loop<K, 5, 5>(a, ar, hA, hB, gs, dt, xs, x, y, z, runA, runB);
    else if (mA == 1 && mB == 1) sigg_loop<K, 1, 1>(a, ar, hA, hB, gs, dt, xs, x, y, z, runA, runB);
    else sigg_loop<K, 6, 0>(a, ar, hA, hB, gs, dt, xs, x, y, z, runA, runB);
}

template __global__ void k_sigg_sparse<6>(SigvArgs);

}  // namespace ae
