// HIP kernels of the statevector engine (gfx950).  See kernels.hpp.
#include <cfloat>

#include "kernels.hpp"

namespace ae {

__global__ void __launch_bounds__(kBlock) k_fill(double *__restrict__ p, long long n, double v)
{
    for (long long i = blockIdx.x * (long long)kBlock + threadIdx.x; i < n; i += (long long)gridDim.x * kBlock)
        p[i] = v;
}

__global__ void __launch_bounds__(kBlock)
    k_absdiff_max(const double *__restrict__ a, const double *__restrict__ ref, long long n, double *__restrict__ out)
{
    __shared__ double sd[kBlock], sr[kBlock];
    double d = 0, r = 0;
    for (long long i = blockIdx.x * (long long)kBlock + threadIdx.x; i < n; i += (long long)gridDim.x * kBlock) {
        const double x = fabs(a[i] - ref[i]);
        d = x <= DBL_MAX ? fmax(d, x) : INFINITY;  // NaN and inf both count as infinitely far
        r = fmax(r, fabs(ref[i]));
    }
    sd[threadIdx.x] = d;
    sr[threadIdx.x] = r;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            sd[threadIdx.x] = fmax(sd[threadIdx.x], sd[threadIdx.x + w]);
            sr[threadIdx.x] = fmax(sr[threadIdx.x], sr[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[2 * blockIdx.x] = sd[0];
        out[2 * blockIdx.x + 1] = sr[0];
    }
}

__global__ void __launch_bounds__(kBlock) k_pattern(double *__restrict__ p, long long n)
{
    for (long long i = blockIdx.x * (long long)kBlock + threadIdx.x; i < n; i += (long long)gridDim.x * kBlock)
        p[i] = (double)((unsigned)(i * 2654435761ull) % 2048u) * (1.0 / 1024.0) - 1.0;
}

__global__ void __launch_bounds__(kBlock) k_copy(const double *__restrict__ src, double *__restrict__ dst, long long n)
{
    for (long long i = blockIdx.x * (long long)kBlock + threadIdx.x; i < n; i += (long long)gridDim.x * kBlock)
        dst[i] = src[i];
}

__global__ void __launch_bounds__(kBlock)
    k_bcast(const double *__restrict__ src, double *__restrict__ dst, long long n, int nvec)
{
    const long long total = n * nvec;
    for (long long t = blockIdx.x * (long long)kBlock + threadIdx.x; t < total; t += (long long)gridDim.x * kBlock)
        dst[t] = src[t % n];
}

// ------------------------------------------------------------- rotations
// Rotation on a pair (b, b') with phase s (src/fusedevolve.cpp:213-234):
//   x_b <- C x_b + s S x_b',   x_b' <- -s S x_b + C x_b'.
template <bool BR>
__global__ void __launch_bounds__(kBlock)
    k_op_generic(double *__restrict__ v0, double *__restrict__ v1, long long vstride, long long N, DevSector sec,
                 const DevOp *__restrict__ ops, int k, const double2 *__restrict__ cs, long long csStride,
                 double *__restrict__ partial, int slots, long long pstride)
{
    __shared__ double red[kBlock / 64];
    const int vb = blockIdx.y;
    v0 += vb * vstride;
    if (BR) v1 += vb * vstride;
    const DevOp o = ops[k];
    const double2 t = cs[vb * csStride + k];
    const double C = t.x, S = t.y;
    double acc = 0;
    for (long long i = blockIdx.x * (long long)kBlock + threadIdx.x; i < N; i += (long long)gridDim.x * kBlock) {
        uint32_t b2;
        const uint32_t b = sec.states[i];
        const int s = pair_sign(o, b, b2);
        if (s == 0 || b2 < b) continue;
        const long long j = sec.rank(b2);
        if (j < 0) continue;
        const double x = v0[i], y = v0[j];
        if (BR) {
            const double lx = v1[i], ly = v1[j];
            acc += s * (lx * y - ly * x);
            v1[i] = s * ly * S + lx * C;
            v1[j] = -s * lx * S + ly * C;
        }
        v0[i] = s * y * S + x * C;
        v0[j] = -s * x * S + y * C;
    }
    if (BR) {
        const double tot = block_sum(acc, red);
        if (threadIdx.x == 0) partial[vb * pstride + (long long)k * slots + blockIdx.x] = tot;
    }
}

template __global__ void k_op_generic<false>(double *, double *, long long, long long, DevSector, const DevOp *, int,
                                             const double2 *, long long, double *, int, long long);
template __global__ void k_op_generic<true>(double *, double *, long long, long long, DevSector, const DevOp *, int,
                                            const double2 *, long long, double *, int, long long);

template <bool BR>
__global__ void __launch_bounds__(kBlock)
    k_phase(double *__restrict__ v0, double *__restrict__ v1, long long vstride, long long N,
            const uint32_t *__restrict__ states, const DevOp *__restrict__ ops, int k0, int k1,
            const double2 *__restrict__ cs, long long csStride, double *__restrict__ partial, int slots,
            long long pstride)
{
    __shared__ double red[kBlock / 64];
    __shared__ DevOp sop[kPhaseMaxOps];
    __shared__ double2 scs[kPhaseMaxOps];
    const int c = blockIdx.y, R = k1 - k0;
    double *re = v0 + 2 * c * vstride, *im = re + vstride;
    double *lre = BR ? v1 + 2 * c * vstride : nullptr, *lim = BR ? lre + vstride : nullptr;
    if ((int)threadIdx.x < R) {
        sop[threadIdx.x] = ops[k0 + threadIdx.x];
        scs[threadIdx.x] = cs[2 * c * csStride + k0 + threadIdx.x];
    }
    __syncthreads();
    double acc[kPhaseMaxOps];
#pragma unroll
    for (int r = 0; r < kPhaseMaxOps; r++) acc[r] = 0;
    for (long long i = blockIdx.x * (long long)kBlock + threadIdx.x; i < N; i += (long long)gridDim.x * kBlock) {
        const uint32_t b = states[i];
        double x = re[i], y = im[i];
        double lx = 0, ly = 0, br = 0;
        if (BR) {
            lx = lre[i];
            ly = lim[i];
            br = ly * x - lx * y;
        }
        bool any = false;
#pragma unroll
        for (int r = 0; r < kPhaseMaxOps; r++) {
            if (r >= R || !number_op_active(sop[r], b)) continue;
            any = true;
            const double C = scs[r].x, S = scs[r].y;
            const double nx = C * x - S * y, ny = S * x + C * y;
            x = nx;
            y = ny;
            if (BR) {
                acc[r] += br;
                const double mx = C * lx - S * ly, my = S * lx + C * ly;
                lx = mx;
                ly = my;
            }
        }
        if (!any) continue;
        re[i] = x;
        im[i] = y;
        if (BR) {
            lre[i] = lx;
            lim[i] = ly;
        }
    }
    if (BR) {
        for (int r = 0; r < R; r++) {
            const double tot = block_sum(acc[r], red);
            if (threadIdx.x == 0) partial[2 * c * pstride + (long long)(k0 + r) * slots + blockIdx.x] = tot;
        }
    }
}

template __global__ void k_phase<false>(double *, double *, long long, long long, const uint32_t *, const DevOp *,
                                        int, int, const double2 *, long long, double *, int, long long);
template __global__ void k_phase<true>(double *, double *, long long, long long, const uint32_t *, const DevOp *,
                                       int, int, const double2 *, long long, double *, int, long long);

// One workgroup per tile: the tile's rows of NV vectors are staged in LDS,
// every operator of the segment is applied there (barrier between
// operators), brackets are reduced per operator, and the rows are written
// back once.  HBM/L2 traffic per segment: one read and one write per vector.
template <int NV, bool BR>
__global__ void __launch_bounds__(kBlock)
    k_seg(SegArgs a, double *__restrict__ v0, double *__restrict__ v1, long long vstride,
          const double2 *__restrict__ cs, long long csStride, double *__restrict__ partial, int slots,
          long long pstride)
{
    extern __shared__ double lds[];
    __shared__ double red[kBlock / 64];
    __shared__ int rowRank[kMaxTileRows];
    __shared__ uint32_t rowAlpha[kMaxTileRows];

    const int tile = blockIdx.x, vb = blockIdx.y;
    v0 += vb * vstride;
    if (NV > 1) v1 += vb * vstride;
    cs += vb * csStride;
    const int r0 = a.tileStart[tile];
    const int R = a.tileStart[tile + 1] - r0;
    const int nb = a.nb;
    if (threadIdx.x < R) {
        const int ra = a.rows[r0 + threadIdx.x];
        rowRank[threadIdx.x] = ra;
        rowAlpha[threadIdx.x] = a.unA[ra];
    }
    __syncthreads();
    double *L0 = lds;
    double *L1 = lds + (size_t)a.maxRows * nb;
    for (int r = 0; r < R; r++) {
        const double *s0 = v0 + (long long)rowRank[r] * nb;
        for (int c = threadIdx.x; c < nb; c += kBlock) L0[r * nb + c] = s0[c];
        if (NV > 1) {
            const double *s1 = v1 + (long long)rowRank[r] * nb;
            for (int c = threadIdx.x; c < nb; c += kBlock) L1[r * nb + c] = s1[c];
        }
    }
    __syncthreads();

    for (int kk = 0; kk < a.k1 - a.k0; kk++) {
        const int k = a.reverse ? a.k1 - 1 - kk : a.k0 + kk;
        const DevOp o = a.ops[k];
        const double2 t = cs[k];
        const double C = t.x, S = t.y;
        double acc = 0;
        for (int r = 0; r < R; r++) {
            const uint32_t alpha = rowAlpha[r];
            const uint32_t ab = alpha << a.n;
            for (int c = threadIdx.x; c < nb; c += kBlock) {
                uint32_t b2;
                const uint32_t b = ab | a.unB[c];
                const int s = pair_sign(o, b, b2);
                if (s == 0 || b2 < b) continue;
                const uint32_t alpha2 = b2 >> a.n;
                const int r2 = (alpha2 == alpha) ? r : a.localRow[a.rkA[alpha2]];
                const int c2 = a.rkB[b2 & a.maskB];
                const int i1 = r * nb + c, i2 = r2 * nb + c2;
                const double x = L0[i1], y = L0[i2];
                if (BR) {
                    const double lx = L1[i1], ly = L1[i2];
                    acc += s * (lx * y - ly * x);
                    L1[i1] = s * ly * S + lx * C;
                    L1[i2] = -s * lx * S + ly * C;
                } else if (NV > 1) {
                    const double lx = L1[i1], ly = L1[i2];
                    L1[i1] = s * ly * S + lx * C;
                    L1[i2] = -s * lx * S + ly * C;
                }
                L0[i1] = s * y * S + x * C;
                L0[i2] = -s * x * S + y * C;
            }
        }
        if (BR) {
            const double tot = block_sum(acc, red);  // contains the barrier
            if (threadIdx.x == 0) partial[vb * pstride + (long long)k * slots + tile] = tot;
        } else {
            __syncthreads();
        }
    }

    for (int r = 0; r < R; r++) {
        double *d0 = v0 + (long long)rowRank[r] * nb;
        for (int c = threadIdx.x; c < nb; c += kBlock) d0[c] = L0[r * nb + c];
        if (NV > 1) {
            double *d1 = v1 + (long long)rowRank[r] * nb;
            for (int c = threadIdx.x; c < nb; c += kBlock) d1[c] = L1[r * nb + c];
        }
    }
}

// Orbit segment: thread per orbit (alpha class x beta class, <= 4 members),
// every operator of the segment applied in registers.  Each unordered member
// pair (p, q) is tested once per operator; the rotation is orientation-free
// (the phase of the partner is minus the phase of the member), so the
// member with the lower slot is taken as "first".  Brackets accumulate per
// operator in registers and are reduced once per block at the end.
template <bool BR>
__global__ void __launch_bounds__(kBlock)
    k_orb(OrbArgs a, double *__restrict__ v0, double *__restrict__ v1, long long vstride,
          const double2 *__restrict__ cs, long long csStride, double *__restrict__ partial, int slots,
          long long pstride)
{
    __shared__ double red[kBlock / 64];
    const int vb = blockIdx.y;
    v0 += vb * vstride;
    if (BR) v1 += vb * vstride;
    cs += vb * csStride;
    const int nk = a.k1 - a.k0;
    double acc[kOrbMaxOps];
#pragma unroll
    for (int t = 0; t < kOrbMaxOps; t++) acc[t] = 0;
    const long long total = (long long)a.nAc * a.nBc;
    for (long long t = blockIdx.x * (long long)kBlock + threadIdx.x; t < total; t += (long long)gridDim.x * kBlock) {
        const int ia = (int)(t / a.nBc), ib = (int)(t - (long long)ia * a.nBc);
        const int4 da = a.aDesc[ia], db = a.bDesc[ib];
        const int Ra = da.y >= 0 ? 2 : 1, Rb = db.y >= 0 ? 2 : 1;
        if (Ra * Rb == 1) continue;  // a lone determinant: no operator of the segment acts on it
        const int ra[2] = {da.x, da.y >= 0 ? da.y : da.x};
        const int rb[2] = {db.x, db.y >= 0 ? db.y : db.x};
        const uint32_t sa[2] = {(uint32_t)da.z << a.n, (uint32_t)da.w << a.n};
        const uint32_t sb[2] = {(uint32_t)db.z, (uint32_t)db.w};
        uint32_t st[4];
        bool ok[4];
        long long idx[4];
        double x[4], l[4];
#pragma unroll
        for (int p = 0; p < 4; p++) {
            const int i = p >> 1, j = p & 1;
            ok[p] = (i < Ra) && (j < Rb);
            st[p] = sa[i] | sb[j];
            idx[p] = (long long)ra[i] * a.nb + rb[j];
            x[p] = ok[p] ? v0[idx[p]] : 0.0;
            if (BR) l[p] = ok[p] ? v1[idx[p]] : 0.0;
        }
#pragma unroll
        for (int kk = 0; kk < kOrbMaxOps; kk++) {
            if (kk >= nk) break;
            const int k = a.reverse ? a.k1 - 1 - kk : a.k0 + kk;
            const DevOp o = a.ops[k];
            const double2 tc = cs[k];
            const double C = tc.x, S = tc.y;
            double br = 0;
#pragma unroll
            for (int p = 0; p < 3; p++) {
                uint32_t b2;
                const int s = ok[p] ? pair_sign(o, st[p], b2) : 0;
#pragma unroll
                for (int q = p + 1; q < 4; q++) {
                    if (s == 0 || !ok[q] || b2 != st[q]) continue;
                    const double xp = x[p], xq = x[q];
                    if (BR) {
                        const double lp = l[p], lq = l[q];
                        br += s * (lp * xq - lq * xp);
                        l[p] = s * lq * S + lp * C;
                        l[q] = -s * lp * S + lq * C;
                    }
                    x[p] = s * xq * S + xp * C;
                    x[q] = -s * xp * S + xq * C;
                }
            }
            acc[kk] += br;
        }
#pragma unroll
        for (int p = 0; p < 4; p++)
            if (ok[p]) {
                v0[idx[p]] = x[p];
                if (BR) v1[idx[p]] = l[p];
            }
    }
    if (BR) {
#pragma unroll
        for (int kk = 0; kk < kOrbMaxOps; kk++) {
            if (kk >= nk) break;
            const int k = a.reverse ? a.k1 - 1 - kk : a.k0 + kk;
            const double tot = block_sum(acc[kk], red);
            if (threadIdx.x == 0) partial[vb * pstride + (long long)k * slots + blockIdx.x] = tot;
        }
    }
}

template <bool BR>
__global__ void __launch_bounds__(kBlock)
    k_orbl(OrbArgs a, double *__restrict__ v0, double *__restrict__ v1, long long vstride,
           const double2 *__restrict__ cs, long long csStride, double *__restrict__ partial, int slots,
           long long pstride)
{
    extern __shared__ double lds[];
    __shared__ double red[kBlock / 64];
    const int ia = blockIdx.x, vb = blockIdx.y, nb = a.nb;
    v0 += vb * vstride;
    if (BR) v1 += vb * vstride;
    cs += vb * csStride;
    const int nk = a.k1 - a.k0;
    const int a0 = a.aStart[ia], Ra = a.aStart[ia + 1] - a0;
    int ra[2];
    ra[0] = a.aRows[a0];
    ra[1] = Ra > 1 ? a.aRows[a0 + 1] : ra[0];
    const uint32_t al0 = a.unA[ra[0]] << a.n, al1 = a.unA[ra[1]] << a.n;
    double *X = lds;           // [2][nb]
    double *L = lds + 2 * nb;  // [2][nb] (adjoint sweep)
    for (int r = 0; r < Ra; r++) {
        const double *s0 = v0 + (long long)ra[r] * nb;
        for (int c = threadIdx.x; c < nb; c += kBlock) X[r * nb + c] = s0[c];
        if (BR) {
            const double *s1 = v1 + (long long)ra[r] * nb;
            for (int c = threadIdx.x; c < nb; c += kBlock) L[r * nb + c] = s1[c];
        }
    }
    __syncthreads();
    double acc[kOrbMaxOps];
#pragma unroll
    for (int t = 0; t < kOrbMaxOps; t++) acc[t] = 0;
    for (int ib = threadIdx.x; ib < a.nBc; ib += kBlock) {
        const int b0 = a.bStart[ib], Rb = a.bStart[ib + 1] - b0;
        if (Ra * Rb == 1) continue;  // a lone determinant: no operator of the segment acts on it
        int rb[2];
        rb[0] = a.bRows[b0];
        rb[1] = Rb > 1 ? a.bRows[b0 + 1] : rb[0];
        const uint32_t be0 = a.unB[rb[0]], be1 = a.unB[rb[1]];
        uint32_t st[4];
        bool ok[4];
        int li[4];
        double x[4], l[4];
#pragma unroll
        for (int p = 0; p < 4; p++) {
            const int i = p >> 1, j = p & 1;
            ok[p] = (i < Ra) && (j < Rb);
            st[p] = (i ? al1 : al0) | (j ? be1 : be0);
            li[p] = i * nb + rb[j];
            x[p] = ok[p] ? X[li[p]] : 0.0;
            if (BR) l[p] = ok[p] ? L[li[p]] : 0.0;
        }
#pragma unroll
        for (int kk = 0; kk < kOrbMaxOps; kk++) {
            if (kk >= nk) break;
            const int k = a.reverse ? a.k1 - 1 - kk : a.k0 + kk;
            const DevOp o = a.ops[k];
            const double2 tc = cs[k];
            const double C = tc.x, S = tc.y;
            double br = 0;
#pragma unroll
            for (int p = 0; p < 3; p++) {
                uint32_t b2;
                const int s = ok[p] ? pair_sign(o, st[p], b2) : 0;
#pragma unroll
                for (int q = p + 1; q < 4; q++) {
                    if (s == 0 || !ok[q] || b2 != st[q]) continue;
                    const double xp = x[p], xq = x[q];
                    if (BR) {
                        const double lp = l[p], lq = l[q];
                        br += s * (lp * xq - lq * xp);
                        l[p] = s * lq * S + lp * C;
                        l[q] = -s * lp * S + lq * C;
                    }
                    x[p] = s * xq * S + xp * C;
                    x[q] = -s * xp * S + xq * C;
                }
            }
            acc[kk] += br;
        }
#pragma unroll
        for (int p = 0; p < 4; p++)
            if (ok[p]) {
                X[li[p]] = x[p];
                if (BR) L[li[p]] = l[p];
            }
    }
    __syncthreads();
    for (int r = 0; r < Ra; r++) {
        double *d0 = v0 + (long long)ra[r] * nb;
        for (int c = threadIdx.x; c < nb; c += kBlock) d0[c] = X[r * nb + c];
        if (BR) {
            double *d1 = v1 + (long long)ra[r] * nb;
            for (int c = threadIdx.x; c < nb; c += kBlock) d1[c] = L[r * nb + c];
        }
    }
    if (BR) {
#pragma unroll
        for (int kk = 0; kk < kOrbMaxOps; kk++) {
            if (kk >= nk) break;
            const int k = a.reverse ? a.k1 - 1 - kk : a.k0 + kk;
            const double tot = block_sum(acc[kk], red);
            if (threadIdx.x == 0) partial[vb * pstride + (long long)k * slots + ia] = tot;
        }
    }
}

template __global__ void k_orbl<false>(OrbArgs, double *, double *, long long, const double2 *, long long, double *,
                                       int, long long);
template __global__ void k_orbl<true>(OrbArgs, double *, double *, long long, const double2 *, long long, double *,
                                      int, long long);

template __global__ void k_orb<false>(OrbArgs, double *, double *, long long, const double2 *, long long, double *, int,
                                      long long);
template __global__ void k_orb<true>(OrbArgs, double *, double *, long long, const double2 *, long long, double *, int,
                                     long long);

// Vector-major [v][N] <-> lane layout, through a 64 x 64 LDS tile.
__global__ void __launch_bounds__(kBlock)
    k_to_lanes(const double *__restrict__ src, double *__restrict__ dst, long long N, int nvec)
{
    __shared__ double tile[kLanes][kLanes + 1];
    const int c = blockIdx.y;
    const long long i0 = blockIdx.x * (long long)kLanes;
    const int tx = threadIdx.x & (kLanes - 1), ty = threadIdx.x / kLanes;
    for (int r = ty; r < kLanes; r += kBlock / kLanes) {  // r = lane (vector), tx = amplitude
        const int v = c * kLanes + r;
        const long long i = i0 + tx;
        tile[r][tx] = (v < nvec && i < N) ? src[(long long)v * N + i] : 0.0;
    }
    __syncthreads();
    for (int r = ty; r < kLanes; r += kBlock / kLanes) {  // r = amplitude, tx = lane
        const long long i = i0 + r;
        if (i < N) dst[((long long)c * N + i) * kLanes + tx] = tile[tx][r];
    }
}

__global__ void __launch_bounds__(kBlock)
    k_from_lanes(const double *__restrict__ src, double *__restrict__ dst, long long N, int nvec)
{
    __shared__ double tile[kLanes][kLanes + 1];
    const int c = blockIdx.y;
    const long long i0 = blockIdx.x * (long long)kLanes;
    const int tx = threadIdx.x & (kLanes - 1), ty = threadIdx.x / kLanes;
    for (int r = ty; r < kLanes; r += kBlock / kLanes) {  // r = amplitude, tx = lane
        const long long i = i0 + r;
        tile[r][tx] = (i < N) ? src[((long long)c * N + i) * kLanes + tx] : 0.0;
    }
    __syncthreads();
    for (int r = ty; r < kLanes; r += kBlock / kLanes) {  // r = lane (vector), tx = amplitude
        const int v = c * kLanes + r;
        const long long i = i0 + tx;
        if (v < nvec && i < N) dst[(long long)v * N + i] = tile[tx][r];
    }
}

__global__ void __launch_bounds__(kBlock)
    k_bcast_lanes(const double *__restrict__ src, double *__restrict__ dst, long long N, int nchunks)
{
    const long long total = N * kLanes * nchunks;
    for (long long t = blockIdx.x * (long long)kBlock + threadIdx.x; t < total; t += (long long)gridDim.x * kBlock)
        dst[t] = src[(t / kLanes) % N];
}

__global__ void __launch_bounds__(kBlock)
    k_reduce_lanes(const double *__restrict__ part, int M, int slots, const int *__restrict__ rows, int nvec,
                   double *__restrict__ out, double scale)
{
    // block per (chunk, operator); thread = lane x slot group
    const int c = blockIdx.x / M, k = blockIdx.x - c * M;
    const int lane = threadIdx.x & (kLanes - 1), g = threadIdx.x / kLanes;
    constexpr int G = kBlock / kLanes;
    __shared__ double red[G][kLanes];
    const double *p = part + ((long long)c * M + k) * slots * kLanes;
    double s = 0;
    const int nrow = rows ? rows[k] : slots;
    for (int j = g; j < nrow; j += G) s += p[(long long)j * kLanes + lane];
    red[g][lane] = s;
    __syncthreads();
    if (g == 0) {
        double tot = 0;
#pragma unroll
        for (int w = 0; w < G; w++) tot += red[w][lane];
        const int v = c * kLanes + lane;
        if (v < nvec) out[(long long)v * M + k] = scale * tot;
    }
}

#define AE_SEG(NV, BR)                                                                                       \
    template __global__ void k_seg<NV, BR>(SegArgs, double *, double *, long long, const double2 *, long long, \
                                           double *, int, long long);
AE_SEG(1, false)
AE_SEG(2, false)
AE_SEG(2, true)
#undef AE_SEG

// ----------------------------------------------------------- Hamiltonian
template <int V>
__global__ void __launch_bounds__(kBlock)
    k_ell(const int32_t *__restrict__ col, const double *__restrict__ val, int W, long long N,
          const double *__restrict__ x, double *__restrict__ y, long long vstride, double *__restrict__ epart,
          int eslots)
{
    __shared__ double red[kBlock / 64];
    const long long j = blockIdx.x * (long long)kBlock + threadIdx.x;
    const int vb = blockIdx.y * V;
    x += vb * vstride;
    y += vb * vstride;
    double acc[V];
#pragma unroll
    for (int t = 0; t < V; t++) acc[t] = 0;
    if (j < N) {
        const int32_t *cp = col + j;
        const double *vp = val + j;
#pragma unroll 4
        for (int e = 0; e < W; e++) {
            const int c = __builtin_nontemporal_load(cp + (size_t)e * N);
            const double v = __builtin_nontemporal_load(vp + (size_t)e * N);
#pragma unroll
            for (int t = 0; t < V; t++) acc[t] += v * x[t * vstride + c];
        }
#pragma unroll
        for (int t = 0; t < V; t++) y[t * vstride + j] = acc[t];
    }
#pragma unroll
    for (int t = 0; t < V; t++) {
        const double p = (j < N) ? x[t * vstride + j] * acc[t] : 0.0;
        const double tot = block_sum(p, red);
        if (threadIdx.x == 0) epart[(long long)(vb + t) * eslots + blockIdx.x] = tot;
    }
}

template __global__ void k_ell<1>(const int32_t *, const double *, int, long long, const double *, double *, long long,
                                  double *, int);
template __global__ void k_ell<2>(const int32_t *, const double *, int, long long, const double *, double *, long long,
                                  double *, int);
template __global__ void k_ell<4>(const int32_t *, const double *, int, long long, const double *, double *, long long,
                                  double *, int);
template __global__ void k_ell<8>(const int32_t *, const double *, int, long long, const double *, double *, long long,
                                  double *, int);

__global__ void __launch_bounds__(kBlock)
    k_rows(const int64_t *__restrict__ rowptr, const int32_t *__restrict__ col, const double *__restrict__ val,
           long long N, const double *__restrict__ x, double *__restrict__ y, long long vstride,
           double *__restrict__ epart, int eslots)
{
    __shared__ double red[kBlock / 64];
    const long long j = blockIdx.x * (long long)kBlock + threadIdx.x;
    const int vb = blockIdx.y;
    x += vb * vstride;
    y += vb * vstride;
    double acc = 0;
    if (j < N) {
        for (int64_t p = rowptr[j]; p < rowptr[j + 1]; p++) acc += x[col[p]] * val[p];
        y[j] = acc;
    }
    const double pe = (j < N) ? x[j] * acc : 0.0;
    const double tot = block_sum(pe, red);
    if (threadIdx.x == 0) epart[(long long)vb * eslots + blockIdx.x] = tot;
}

__global__ void __launch_bounds__(kBlock)
    k_reduce(const double *__restrict__ part, int nrows, int ncols, long long rstride, long long cstride,
             double *__restrict__ out, double scale)
{
    __shared__ double red[kBlock / 64];
    const int r = blockIdx.x;
    if (r >= nrows) return;
    double acc = 0;
    for (int c = threadIdx.x; c < ncols; c += kBlock) acc += part[r * rstride + c * cstride];
    const double tot = block_sum(acc, red);
    if (threadIdx.x == 0) out[r] = scale * tot;
}

// Value of the reference's two-body operator a^+_a a^+_b a_c a_d (a < b,
// c < d) from chemist integrals in the _twoEInts.bin layout, and its
// sign mask (src/hamiltonianmatrix.cpp:445-529).  Values at or below
// 1e-13 in magnitude are dropped as the reference drops them (:483, :509).
__device__ __forceinline__ double two_body(int a, int b, int c, int d, int n, const double *__restrict__ g,
                                           uint32_t &mask)
{
    const int sa = a >= n, sb = b >= n, sc = c >= n, sd = d >= n;
    const int pa = a - sa * n, pb = b - sb * n, pc = c - sc * n, pd = d - sd * n;
    double E = 0;
    if (sa == sc && sb == sd) E += g[(((size_t)pb * n + pd) * n + pa) * n + pc];
    if (sa == sd && sb == sc) E -= g[(((size_t)pb * n + pc) * n + pa) * n + pd];
    const uint32_t m = ((1u << a) - 1) ^ ((1u << b) - 1) ^ ((1u << c) - 1) ^ ((1u << d) - 1);
    mask = m & ~((1u << a) | (1u << b) | (1u << c) | (1u << d));
    return fabs(E) > 1e-13 ? E : 0.0;
}

__device__ __forceinline__ double one_body(int a, int c, int n, const double *__restrict__ h, uint32_t &mask)
{
    const double E = h[(size_t)(a % n) * n + (c % n)];
    mask = (((1u << a) - 1) ^ ((1u << c) - 1)) & ~((1u << a) | (1u << c));
    return fabs(E) > 1e-13 ? E : 0.0;
}

__device__ __forceinline__ double signed_by(uint32_t jb, uint32_t mask, double v)
{
    return (__popc(jb & mask) & 1) ? -v : v;
}

// Row j = (alpha, beta) of y = M^T x.  Entry order (identical for every row):
// diagonal, down singles, up singles, down-down doubles, up-up doubles,
// up-down doubles.  An entry (i, M(i,j)) sums every operator of the
// reference's list that maps basis state i to j (create bits C subset of j,
// destroy bits D subset of i), each with sign (-1)^popc(j & mask)
// (the matrix-free rule, src/hamiltonianmatrix.cpp:915-1009).
__global__ void __launch_bounds__(kBlock) k_ell_build(EllBuildArgs a)
{
    const long long j = blockIdx.x * (long long)kBlock + threadIdx.x;
    if (j >= a.N) return;
    const int n = a.n;
    const long long ra = j / a.nb;
    const int rb = (int)(j - ra * a.nb);
    const uint32_t alpha = a.unA[ra], beta = a.unB[rb];
    const uint32_t jb = (alpha << n) | beta;
    int occ[32], occD[16], virD[16], occU[16], virU[16];
    int no = 0, nod = 0, nvd = 0, nou = 0, nvu = 0;
    for (int p = 0; p < n; p++) {
        if (beta >> p & 1) { occD[nod++] = p; occ[no++] = p; }
        else virD[nvd++] = p;
    }
    for (int p = 0; p < n; p++) {
        if (alpha >> p & 1) { occU[nou++] = p + n; occ[no++] = p + n; }
        else virU[nvu++] = p + n;
    }
    int32_t *col = a.col + j;
    double *val = a.val + j;
    const size_t N = (size_t)a.N;
    int e = 0;
    uint32_t m;
    // diagonal
    {
        double v = 0;
        for (int x = 0; x < no; x++) v += signed_by(jb, 0, one_body(occ[x], occ[x], n, a.h, m));
        for (int x = 0; x < no; x++)
            for (int y = x + 1; y < no; y++) {
                const double t = two_body(occ[x], occ[y], occ[x], occ[y], n, a.g, m);
                v += signed_by(jb, m, t);
            }
        col[e * N] = (int32_t)j;
        val[e * N] = v;
        e++;
    }
    // singles: i = j - {r} + {q}; one-body op (C={r}, D={q}) plus two-body
    // ops (C={r,x}, D={q,x}) for every spectator x occupied in both.
    for (int spin = 0; spin < 2; spin++) {
        const int *oc = spin ? occU : occD;
        const int *vi = spin ? virU : virD;
        const int nocc = spin ? nou : nod, nvir = spin ? nvu : nvd;
        for (int ir = 0; ir < nocc; ir++)
            for (int iq = 0; iq < nvir; iq++) {
                const int r = oc[ir], q = vi[iq];
                double v = 0.0;
                {
                    const double t = one_body(r, q, n, a.h, m);
                    v += signed_by(jb, m, t);
                }
                for (int xx = 0; xx < no; xx++) {
                    const int x = occ[xx];
                    if (x == r) continue;
                    const int c1 = r < x ? r : x, c2 = r < x ? x : r;
                    const int d1 = q < x ? q : x, d2 = q < x ? x : q;
                    const double t = two_body(c1, c2, d1, d2, n, a.g, m);
                    v += signed_by(jb, m, t);
                }
                const uint32_t ib = jb ^ (1u << r) ^ (1u << q);
                const long long ci = (long long)a.rkA[ib >> n] * a.nb + a.rkB[ib & ((1u << n) - 1)];
                col[e * N] = (int32_t)ci;
                val[e * N] = v;
                e++;
            }
    }
    // same-spin doubles: C = {r1 < r2}, D = {q1 < q2}
    for (int spin = 0; spin < 2; spin++) {
        const int *oc = spin ? occU : occD;
        const int *vi = spin ? virU : virD;
        const int nocc = spin ? nou : nod, nvir = spin ? nvu : nvd;
        for (int i1 = 0; i1 < nocc; i1++)
            for (int i2 = i1 + 1; i2 < nocc; i2++)
                for (int j1 = 0; j1 < nvir; j1++)
                    for (int j2 = j1 + 1; j2 < nvir; j2++) {
                        const int r1 = oc[i1], r2 = oc[i2], q1 = vi[j1], q2 = vi[j2];
                        const double t = two_body(r1, r2, q1, q2, n, a.g, m);
                        const uint32_t ib = jb ^ (1u << r1) ^ (1u << r2) ^ (1u << q1) ^ (1u << q2);
                        const long long ci = (long long)a.rkA[ib >> n] * a.nb + a.rkB[ib & ((1u << n) - 1)];
                        col[e * N] = (int32_t)ci;
                        val[e * N] = signed_by(jb, m, t);
                        e++;
                    }
    }
    // opposite-spin doubles: C = {r_down < r_up}, D = {q_down < q_up}
    for (int iu = 0; iu < nou; iu++)
        for (int ju = 0; ju < nvu; ju++)
            for (int id = 0; id < nod; id++)
                for (int jd = 0; jd < nvd; jd++) {
                    const int ru = occU[iu], qu = virU[ju], rd = occD[id], qd = virD[jd];
                    const double t = two_body(rd, ru, qd, qu, n, a.g, m);
                    const uint32_t ib = jb ^ (1u << ru) ^ (1u << qu) ^ (1u << rd) ^ (1u << qd);
                    const long long ci = (long long)a.rkA[ib >> n] * a.nb + a.rkB[ib & ((1u << n) - 1)];
                    col[e * N] = (int32_t)ci;
                    val[e * N] = signed_by(jb, m, t);
                    e++;
                }
}

// ------------------------------------------------------- matrix-free sigma
__device__ __forceinline__ double thr13(double v) { return fabs(v) > 1e-13 ? v : 0.0; }

// (-1)^(number of occupied orbitals strictly between r and q)
__device__ __forceinline__ int between_parity(uint32_t s, int r, int q)
{
    const int lo = r < q ? r : q, hi = r < q ? q : r;
    const uint32_t m = ((1u << hi) - 1) & ~((1u << (lo + 1)) - 1);
    return __popc(s & m) & 1;
}

// Signed single-excitation value of one spin string (orbitals 0..n-1,
// spin-orbital offset `off`): one-body term plus the same-spin spectator
// two-body terms, each with its own sign (src/hamiltonianmatrix.cpp:512-541).
__device__ double same_spin_single(uint32_t s, int off, int r, int q, int n, const double *__restrict__ h,
                                   const double *__restrict__ g)
{
    const uint32_t jb = s << off;
    uint32_t m;
    double v = 0.0;
    {
        const double t = one_body(r + off, q + off, n, h, m);
        v += signed_by(jb, m, t);
    }
    for (uint32_t o = s; o; o &= o - 1) {
        const int x = __ffs(o) - 1;
        if (x == r) continue;
        const int R = r + off, Q = q + off, X = x + off;
        const double t = two_body(R < X ? R : X, R < X ? X : R, Q < X ? Q : X, Q < X ? X : Q, n, g, m);
        v += signed_by(jb, m, t);
    }
    return v;
}

__global__ void __launch_bounds__(kBlock) k_mf_tables(MfArgs a)
{
    const long long idx = blockIdx.x * (long long)kBlock + threadIdx.x;
    const int n = a.n;
    const uint32_t full = (1u << n) - 1;
    if (idx < a.N) {
        const long long ra = idx / a.nb;
        const int rb = (int)(idx - ra * a.nb);
        const uint32_t jb = (a.unA[ra] << n) | a.unB[rb];
        uint32_t m;
        double v = 0;
        for (uint32_t o = jb; o; o &= o - 1) {
            const int x = __ffs(o) - 1;
            v += one_body(x, x, n, a.h, m);
            for (uint32_t o2 = o & (o - 1); o2; o2 &= o2 - 1) {
                const int y = __ffs(o2) - 1;
                v += two_body(x, y, x, y, n, a.g, m);  // mask 0: no sign
            }
        }
        a.diag[idx] = v;
    }
    if (idx < a.nb) {
        const uint32_t s = a.unB[idx];
        int t = 0;
        for (uint32_t o = s; o; o &= o - 1)
            for (uint32_t w = ~s & full; w; w &= w - 1, t++)
                a.FB[(long long)t * a.nb + idx] = same_spin_single(s, 0, __ffs(o) - 1, __ffs(w) - 1, n, a.h, a.g);
    }
    if (idx < a.na) {
        const uint32_t s = a.unA[idx];
        int t = 0;
        for (uint32_t o = s; o; o &= o - 1)
            for (uint32_t w = ~s & full; w; w &= w - 1, t++)
                a.FA[(long long)t * a.na + idx] = same_spin_single(s, n, __ffs(o) - 1, __ffs(w) - 1, n, a.h, a.g);
    }
}

// sigma(alpha, .) = (M^T x)(alpha, .): the row's own beta singles and
// doubles (x row staged in LDS), then for every alpha single the source row
// in LDS with the up-down doubles through that row, then the up-up doubles.
// Up-down double (r_a q_a | r_b q_b): value g[r_a][q_a][r_b][q_b] in the
// reference's lookup, sign = product of the two strings' between-parities.
__global__ void __launch_bounds__(kBlock)
    k_mf_sigma(MfArgs a, const double *__restrict__ x, double *__restrict__ y, long long vstride,
               double *__restrict__ epart)
{
    extern __shared__ double sm[];
    __shared__ double red[kBlock / 64];
    const int n = a.n, nb = a.nb, nn = n * n;
    const uint32_t full = (1u << n) - 1;
    double *xrow = sm;                       // nb
    double *gsl = xrow + nb;                 // n*n
    double *Ja = gsl + nn;                   // n*n
    short *rkb = (short *)(Ja + nn);         // 2^n
    const int ra = blockIdx.x, vb = blockIdx.y;
    x += vb * vstride;
    y += vb * vstride;
    const uint32_t alpha = a.unA[ra];
    for (int s = threadIdx.x; s < (1 << n); s += kBlock) rkb[s] = (short)a.rkB[s];
    for (int e = threadIdx.x; e < nn; e += kBlock) {
        double v = 0;
        for (uint32_t o = alpha; o; o &= o - 1) {
            const int xo = __ffs(o) - 1;
            v += thr13(a.g[(size_t)(xo * n + xo) * nn + e]);
        }
        Ja[e] = v;
    }
    const double *xa = x + (long long)ra * nb;
    for (int c = threadIdx.x; c < nb; c += kBlock) xrow[c] = xa[c];
    __syncthreads();

    constexpr int JM = 16;  // up to 16 * kBlock beta strings (n <= 14)
    double acc[JM];
#pragma unroll
    for (int j = 0; j < JM; j++) acc[j] = 0;

    // same alpha: diagonal, beta singles, beta-beta doubles
#pragma unroll
    for (int j = 0; j < JM; j++) {
        const int c = threadIdx.x + j * kBlock;
        if (c >= nb) break;
        const uint32_t b = a.unB[c];
        double v = a.diag[(long long)ra * nb + c] * xrow[c];
        int t = 0;
        for (uint32_t o = b; o; o &= o - 1) {
            const int r = __ffs(o) - 1;
            for (uint32_t w = ~b & full; w; w &= w - 1, t++) {
                const int q = __ffs(w) - 1;
                const int c2 = rkb[b ^ (1u << r) ^ (1u << q)];
                const double jv = Ja[r * n + q];
                const double val = a.FB[(long long)t * nb + c] + (between_parity(b, r, q) ? -jv : jv);
                v += val * xrow[c2];
            }
        }
        for (uint32_t o1 = b; o1; o1 &= o1 - 1) {
            const int r1 = __ffs(o1) - 1;
            for (uint32_t o2 = o1 & (o1 - 1); o2; o2 &= o2 - 1) {
                const int r2 = __ffs(o2) - 1;
                for (uint32_t w1 = ~b & full; w1; w1 &= w1 - 1) {
                    const int q1 = __ffs(w1) - 1;
                    for (uint32_t w2 = w1 & (w1 - 1); w2; w2 &= w2 - 1) {
                        const int q2 = __ffs(w2) - 1;
                        uint32_t m;
                        const double tv = two_body(r1, r2, q1, q2, n, a.g, m);
                        const int c2 = rkb[b ^ (1u << r1) ^ (1u << r2) ^ (1u << q1) ^ (1u << q2)];
                        v += signed_by(b, m, tv) * xrow[c2];
                    }
                }
            }
        }
        acc[j] = v;
    }

    // alpha singles and up-down doubles, one source row at a time
    int tA = 0;
    for (uint32_t oa = alpha; oa; oa &= oa - 1) {
        const int rA = __ffs(oa) - 1;
        for (uint32_t wa = ~alpha & full; wa; wa &= wa - 1, tA++) {
            const int qA = __ffs(wa) - 1;
            const uint32_t alpha2 = alpha ^ (1u << rA) ^ (1u << qA);
            const int ra2 = a.rkA[alpha2];
            const int sA = between_parity(alpha, rA, qA);
            __syncthreads();
            const double *xs = x + (long long)ra2 * nb;
            for (int c = threadIdx.x; c < nb; c += kBlock) xrow[c] = xs[c];
            const double *gb = a.g + (size_t)(rA * n + qA) * nn;
            for (int e = threadIdx.x; e < nn; e += kBlock) gsl[e] = thr13(gb[e]);
            __syncthreads();
            const double FAv = a.FA[(long long)tA * a.na + ra];
#pragma unroll
            for (int j = 0; j < JM; j++) {
                const int c = threadIdx.x + j * kBlock;
                if (c >= nb) break;
                const uint32_t b = a.unB[c];
                double jb = 0, dsum = 0;
                for (uint32_t o = b; o; o &= o - 1) {
                    const int r = __ffs(o) - 1;
                    jb += gsl[r * n + r];
                    for (uint32_t w = ~b & full; w; w &= w - 1) {
                        const int q = __ffs(w) - 1;
                        const double t = gsl[r * n + q] * xrow[rkb[b ^ (1u << r) ^ (1u << q)]];
                        dsum += between_parity(b, r, q) ? -t : t;
                    }
                }
                const double single = FAv + (sA ? -jb : jb);
                acc[j] += single * xrow[c] + (sA ? -dsum : dsum);
            }
        }
    }

    // up-up doubles: value independent of beta
    const uint32_t jbA = alpha << n;
    for (uint32_t o1 = alpha; o1; o1 &= o1 - 1) {
        const int r1 = __ffs(o1) - 1;
        for (uint32_t o2 = o1 & (o1 - 1); o2; o2 &= o2 - 1) {
            const int r2 = __ffs(o2) - 1;
            for (uint32_t w1 = ~alpha & full; w1; w1 &= w1 - 1) {
                const int q1 = __ffs(w1) - 1;
                for (uint32_t w2 = w1 & (w1 - 1); w2; w2 &= w2 - 1) {
                    const int q2 = __ffs(w2) - 1;
                    uint32_t m;
                    const double tv = two_body(r1 + n, r2 + n, q1 + n, q2 + n, n, a.g, m);
                    const double val = signed_by(jbA, m, tv);
                    const int ra2 = a.rkA[alpha ^ (1u << r1) ^ (1u << r2) ^ (1u << q1) ^ (1u << q2)];
                    const double *xs = x + (long long)ra2 * nb;
#pragma unroll
                    for (int j = 0; j < JM; j++) {
                        const int c = threadIdx.x + j * kBlock;
                        if (c >= nb) break;
                        acc[j] += val * xs[c];
                    }
                }
            }
        }
    }

    double ep = 0;
    double *ya = y + (long long)ra * nb;
#pragma unroll
    for (int j = 0; j < JM; j++) {
        const int c = threadIdx.x + j * kBlock;
        if (c >= nb) break;
        ya[c] = acc[j];
        ep += xa[c] * acc[j];
    }
    const double tot = block_sum(ep, red);
    if (threadIdx.x == 0) epart[(long long)vb * a.na + ra] = tot;
}

// ------------------------------------------- matrix-free sigma, lane layout
// Same-spin single/double parts as dense rows (k_dense_spin): row s of
// V[s][s'] = signed value of every same-spin single (F table) and double
// taking string s' to s.  Used as DGEMM operands for the alpha-alpha and
// the beta-beta (alpha-independent) terms.
__global__ void __launch_bounds__(kBlock) k_dense_spin(MfArgs a, int spin, double *__restrict__ V)
{
    const int s = blockIdx.x * kBlock + threadIdx.x;
    const int cnt = spin ? a.na : a.nb;
    if (s >= cnt) return;
    const int n = a.n, off = spin ? n : 0;
    const uint32_t full = (1u << n) - 1;
    const uint32_t str = spin ? a.unA[s] : a.unB[s];
    const int32_t *rk = spin ? a.rkA : a.rkB;
    const double *F = spin ? a.FA : a.FB;
    double *row = V + (size_t)s * cnt;
    int t = 0;
    for (uint32_t o = str; o; o &= o - 1)
        for (uint32_t w = ~str & full; w; w &= w - 1, t++) {
            const int r = __ffs(o) - 1, q = __ffs(w) - 1;
            row[rk[str ^ (1u << r) ^ (1u << q)]] += F[(size_t)t * cnt + s];
        }
    const uint32_t js = str << off;
    for (uint32_t o1 = str; o1; o1 &= o1 - 1) {
        const int r1 = __ffs(o1) - 1;
        for (uint32_t o2 = o1 & (o1 - 1); o2; o2 &= o2 - 1) {
            const int r2 = __ffs(o2) - 1;
            for (uint32_t w1 = ~str & full; w1; w1 &= w1 - 1) {
                const int q1 = __ffs(w1) - 1;
                for (uint32_t w2 = w1 & (w1 - 1); w2; w2 &= w2 - 1) {
                    const int q2 = __ffs(w2) - 1;
                    uint32_t m;
                    const double tv = two_body(r1 + off, r2 + off, q1 + off, q2 + off, n, a.g, m);
                    row[rk[str ^ (1u << r1) ^ (1u << r2) ^ (1u << q1) ^ (1u << q2)]] += signed_by(js, m, tv);
                }
            }
        }
    }
}

// Beta entry table of the matrix-free sigma: per beta string b, its singles
// (occupied r ascending x empty q ascending) and then its diagonal entries
// (r = q over occupied r), padded to ntp entries with an entry whose
// coefficient slot is zero.  With sepdiag the diagonal entries are left out
// and the last slot holds the string b.  Entry = (byte offset J * xunit of the source
// beta J in a staged row) | (byte offset of the coefficient
// (e + parity * nnoff), e = r * n + q, in a coefficient row) << cshift.
__global__ void __launch_bounds__(kBlock)
    k_beta_table(MfArgs a, uint32_t *__restrict__ ts, int ntp, int xunit, int nnoff, int cshift, int sepdiag)
{
    const int b = blockIdx.x * kBlock + threadIdx.x;
    if (b >= a.nb) return;
    const int n = a.n;
    const uint32_t full = (1u << n) - 1, s = a.unB[b];
    uint32_t *row = ts + (size_t)b * ntp;
    int t = 0;
    for (uint32_t o = s; o; o &= o - 1)
        for (uint32_t w = ~s & full; w; w &= w - 1) {
            const int r = __ffs(o) - 1, q = __ffs(w) - 1;
            const int J = a.rkB[s ^ (1u << r) ^ (1u << q)];
            const int e = r * n + q + (between_parity(s, r, q) ? nnoff : 0);
            row[t++] = (uint32_t)(J * xunit) | ((uint32_t)(e * 8) << cshift);
        }
    if (!sepdiag)
        for (uint32_t o = s; o; o &= o - 1) {
            const int r = __ffs(o) - 1;
            row[t++] = (uint32_t)(b * xunit) | ((uint32_t)((r * n + r) * 8) << cshift);
        }
    for (; t < ntp; t++) row[t] = (uint32_t)(2 * nnoff * 8) << cshift;
    if (sepdiag) row[ntp - 1] = s;  // the string itself, for the diagonal sums
}

// Lane-layout sigma, sparse part: for output (alpha, beta), the alpha-beta
// doubles and the beta-diagonal part of the alpha singles through every alpha
// single's source row, the alpha-dependent spectator part of the beta singles
// (identity row), and the diagonal.  16 vectors per workgroup, one staged
// source row (nb x 16 doubles), one LDS read per entry.  A wave holds four target betas, one
// per 16-lane DPP row (lane = vector).  Each lane loads the table entry of
// its row's target with index (lane mod 16) and that entry's coefficient;
// the 16 entries of a chunk are then broadcast across the row with
// row_newbcast (the source offset folds into the address add), so every
// entry costs one conflict-free 128-byte row read of the staged source and
// one FMA.  The next source row is fetched into registers while the current
// one is summed.
typedef __attribute__((address_space(3))) double lds_f64;
__device__ __forceinline__ double lds_at(uint32_t byteaddr) { return *(const lds_f64 *)(size_t)byteaddr; }
template <int J>
__device__ __forceinline__ double row_bcast(double v)
{
    return __builtin_amdgcn_update_dpp(0.0, v, 0x150 + J, 0xf, 0xf, true);
}
template <int J>
__device__ __forceinline__ uint32_t row_bcast(uint32_t v)
{
    return __builtin_amdgcn_update_dpp(0u, v, 0x150 + J, 0xf, 0xf, true);
}

// acc += (c of lane J of this 16-lane row) * x: the broadcast rides on the
// FMA (DPP row_newbcast on a 64-bit VALU op); c comes from an LDS read, so no
// VALU result feeds the DPP operand
template <int J>
__device__ __forceinline__ void fma_row_bcast(double &acc, double c, double x)
{
    asm("v_fmac_f64_dpp %0, %1, %2 row_newbcast:%3 row_mask:0xf bank_mask:0xf bound_ctrl:1"
        : "+v"(acc)
        : "v"(c), "v"(x), "i"(J));
}

// four entries of a 16-entry chunk (J, J+1, J+2, J+3): four partial sums,
// so four row reads can be in flight per wave
template <int J>
__device__ __forceinline__ void sigw_group(double cl, uint32_t xb, uint32_t lb, double (&s)[4])
{
    // four row reads issued before the four FMAs that use them
    const double x0 = lds_at(row_bcast<J>(xb) + lb), x1 = lds_at(row_bcast<J + 1>(xb) + lb);
    const double x2 = lds_at(row_bcast<J + 2>(xb) + lb), x3 = lds_at(row_bcast<J + 3>(xb) + lb);
    __builtin_amdgcn_sched_barrier(0);  // keep the four reads ahead of the FMAs
    fma_row_bcast<J>(s[0], cl, x0);
    fma_row_bcast<J + 1>(s[1], cl, x1);
    fma_row_bcast<J + 2>(s[2], cl, x2);
    fma_row_bcast<J + 3>(s[3], cl, x3);
}

// one chunk of 16 table entries, the first min(G, 4) groups of 4
template <int G>
__device__ __forceinline__ void sigw_chunk(uint32_t gsb, uint32_t ent, uint32_t lb, double (&s)[4])
{
    const double cl = lds_at(gsb + (ent >> 17));
    const uint32_t xb = ent & 0x1ffffu;
    sigw_group<0>(cl, xb, lb, s);
    if constexpr (G > 1) sigw_group<4>(cl, xb, lb, s);
    if constexpr (G > 2) sigw_group<8>(cl, xb, lb, s);
    if constexpr (G > 3) sigw_group<12>(cl, xb, lb, s);
}

template <int NG>
__global__ void __launch_bounds__(kSigWThreads) k_sigw_sparse(SigvArgs a)
{
    extern __shared__ double sm[];
    const MfArgs &m = a.m;
    const int n = m.n, nb = m.nb, nn = n * n;
    // XCD-aware order: the 4 sub-chunks of one alpha row run on one XCD
    const int bid = blockIdx.x, xcd = bid & 7, k = bid >> 3;
    const int ar = xcd + 8 * (k >> 2), sub = k & 3;
    if (ar >= m.na) return;
    const long long coff = (long long)blockIdx.y * a.cstride + sub * kSigWVC;
    const double *__restrict__ x = a.x + coff;
    double *__restrict__ y = a.y + coff;
    double *gs = sm, *dt = sm + kSigGS, *xs = dt + 128;
    const int tid = threadIdx.x, lane = tid & (kLanes - 1), w = tid / kLanes;
    const int v = lane & 15, trow = w * 4 + (lane >> 4);
    const int wu = __builtin_amdgcn_readfirstlane(w);
    const uint32_t alpha = m.unA[ar], full = (1u << n) - 1;
    const int ka = __popc(alpha), ne = n - ka;
    const int R = 1 + ka * ne;
    const uint32_t gsb = (uint32_t)(size_t)(lds_f64 *)gs;
    const uint32_t lb = (uint32_t)(size_t)(lds_f64 *)xs + v * 8;
    constexpr uint32_t kPad = (uint32_t)(2 * kSigNN * 8) << 17;  // zero coefficient, source beta 0
    const uint32_t toff = (uint32_t)(trow * kSigWNtp + v);
    double acc[kSigWNI];
#pragma unroll
    for (int i = 0; i < kSigWNI; i++) acc[i] = 0;
    for (int r = 0; r < R; r++) {
        // source row: r = 0 the identity row (J = ar), else the alpha single
        // tA = r - 1 (occupied rA, empty qA by rank) with its coefficients
        int J = ar, rA = 0, qA = 0;
        if (r > 0) {
            const int tA = r - 1, ir = tA / ne, iq = tA - ir * ne;
            uint32_t oo = alpha, ww = ~alpha & full;
            for (int i = 0; i < ir; i++) oo &= oo - 1;
            for (int i = 0; i < iq; i++) ww &= ww - 1;
            rA = __ffs(oo) - 1;
            qA = __ffs(ww) - 1;
            J = m.rkA[alpha ^ (1u << rA) ^ (1u << qA)];
        }
        __syncthreads();  // every read of the previous row is done
        {
            // the source row into LDS by asynchronous global-to-LDS copies, all
            // issued before any wait: 8 betas (1 KB) per wave instruction,
            // lane l -> beta 8c + l/8, vectors 2 (l mod 8) and 2 (l mod 8) + 1
            // (a load/store loop waited for each load in turn)
            const double *src = x + (long long)J * nb * kLanes + (lane & 7) * 2;
            for (int c = wu; c * 8 < nb; c += kSigWThreads / kLanes) {
                const int beta = c * 8 + (lane >> 3);
                __builtin_amdgcn_global_load_lds((const void *)(src + (long long)(beta < nb ? beta : 0) * kLanes),
                                                 (__attribute__((address_space(3))) void *)(xs + c * 128), 16, 0,
                                                 0);
            }
        }
        if (r == 0) {
            // identity row: J_alpha[e] = sum over occupied alpha o of g[oo][e] (r != q only)
            for (int e = tid; e < nn; e += kSigWThreads) {
                double vj = 0;
                if (e / n != e % n)
                    for (uint32_t o = alpha; o; o &= o - 1) {
                        const int xo = __ffs(o) - 1;
                        vj += thr13(m.g[(size_t)(xo * n + xo) * nn + e]);
                    }
                gs[e] = vj;
                gs[kSigNN + e] = -vj;
            }
        } else {
            const bool neg = between_parity(alpha, rA, qA);
            const double *gb = m.g + (size_t)(rA * n + qA) * nn;
            for (int e = tid; e < nn; e += kSigWThreads) {
                const double cf = neg ? -thr13(gb[e]) : thr13(gb[e]);
                gs[e] = cf;
                gs[kSigNN + e] = -cf;
            }
        }
        if (tid < kSigGS - 2 * kSigNN) gs[2 * kSigNN + tid] = 0.0;
        // beta-diagonal sums over the occupied orbitals of a string, by its
        // low and high 6 bits: dt[t] (bits 0-5), dt[64 + t] (bits 6-11); zero
        // for the identity row, whose diagonal is the diag array
        if (tid < 128) {
            const int hi = tid >> 6, t = tid & 63;
            double d = 0;
            if (r > 0) {
                const bool neg = between_parity(alpha, rA, qA);
                const double *gb = m.g + (size_t)(rA * n + qA) * nn;
                // the six loads first, then the sums (a conditional load per
                // term waited for each in turn)
                double gv[6];
#pragma unroll
                for (int b = 0; b < 6; b++) {
                    const int q = b + 6 * hi;
                    gv[b] = gb[(q < n ? q : 0) * (n + 1)];
                }
#pragma unroll
                for (int b = 0; b < 6; b++)
                    if (b + 6 * hi < n && ((t >> b) & 1)) d += neg ? -thr13(gv[b]) : thr13(gv[b]);
            }
            dt[tid] = d;
        }
        __builtin_amdgcn_s_waitcnt(0x0f70);  // vmcnt(0): this thread's row copies have landed
        __syncthreads();
        // table entries of the next pass are loaded while this pass sums
        // (32-bit element offsets from the table base); the empty asm keeps
        // the compiler from hoisting every pass's entries out of the row
        // loop, which would hold 45 more registers
        uint32_t tof = toff;
        asm volatile("" : "+v"(tof));
        uint32_t ent[kSigWH], entn[kSigWH];
#pragma unroll
        for (int h = 0; h < kSigWH; h++) ent[h] = trow < nb ? a.ts[tof + 16 * h] : kPad;
#pragma unroll
        for (int i = 0; i < kSigWNI; i++) {
            // no target of this wave in this pass (or later ones): the partial
            // last pass leaves its LDS bandwidth to the waves that have some
            if (64 * i + 4 * wu >= nb) break;
            const int tgt = trow + 64 * i, tn = tgt + 64;
#pragma unroll
            for (int h = 0; h < kSigWH; h++)
                entn[h] = tn < nb ? a.ts[tof + (uint32_t)(64 * kSigWNtp * (i + 1) + 16 * h)] : kPad;
            double s[4] = {0, 0, 0, 0};
            // the singles: NG groups of 4 entries in chunks of 16
            sigw_chunk<NG>(gsb, ent[0], lb, s);
            if constexpr (NG > 4) sigw_chunk<NG - 4>(gsb, ent[1], lb, s);
            if constexpr (NG > 8) sigw_chunk<NG - 8>(gsb, ent[2], lb, s);
            acc[i] += (s[0] + s[1]) + (s[2] + s[3]);
            // beta diagonal: the sums over the target string's occupied
            // orbitals (the string is the last table slot)
            const uint32_t bs = row_bcast<15>(ent[kSigWH - 1]);
            if (tgt < nb) acc[i] += (dt[bs & 63] + dt[64 + ((bs >> 6) & 63)]) * xs[tgt * kSigWVC + v];
            if (r == 0 && tgt < nb) acc[i] += m.diag[(long long)ar * nb + tgt] * xs[tgt * kSigWVC + v];
#pragma unroll
            for (int h = 0; h < kSigWH; h++) ent[h] = entn[h];
        }
    }
#pragma unroll
    for (int i = 0; i < kSigWNI; i++) {
        const int tgt = trow + 64 * i;
        if (tgt < nb) y[((long long)ar * nb + tgt) * kLanes + v] = acc[i];
    }
}

// Lane-layout sigma, sparse part, with the source row split in two halves
// by the occupation of beta orbital kSigHOrb and double-buffered: while one
// half is summed, the next half is copied from HBM into the other LDS buffer
// by asynchronous global-to-LDS loads (no registers held).  A single with
// r, q != kSigHOrb keeps the half, one moving that orbital crosses it, so for
// n = 2 kd every target has S same-half and X cross-half entries; targets are
// ordered by half (each half padded to a multiple of 4), so a wave's four
// targets share their counts.  Per phase a target sums its entries for the
// staged half; its diagonal terms come in the phase of its own half.
template <int J>
__device__ __forceinline__ void sigw_pair(double cl, uint32_t xb, uint32_t lb, double (&s)[4])
{
    const double x0 = lds_at(row_bcast<J>(xb) + lb), x1 = lds_at(row_bcast<J + 1>(xb) + lb);
    __builtin_amdgcn_sched_barrier(0);
    fma_row_bcast<J>(s[0], cl, x0);
    fma_row_bcast<J + 1>(s[1], cl, x1);
}
template <int J>
__device__ __forceinline__ void sigw_single(double cl, uint32_t xb, uint32_t lb, double (&s)[4])
{
    const double x0 = lds_at(row_bcast<J>(xb) + lb);
    fma_row_bcast<J>(s[2], cl, x0);
}
// the first NIN (1..16) entries of a chunk
template <int NIN>
__device__ __forceinline__ void sigh_chunk(uint32_t gsb, uint32_t ent, uint32_t lb, double (&s)[4])
{
    const double cl = lds_at(gsb + (ent >> 17));
    const uint32_t xb = ent & 0x1ffffu;
    if constexpr (NIN >= 4) sigw_group<0>(cl, xb, lb, s);
    if constexpr (NIN >= 8) sigw_group<4>(cl, xb, lb, s);
    if constexpr (NIN >= 12) sigw_group<8>(cl, xb, lb, s);
    if constexpr (NIN >= 16) sigw_group<12>(cl, xb, lb, s);
    constexpr int G = NIN & ~3;
    if constexpr ((NIN & 3) >= 2) sigw_pair<G>(cl, xb, lb, s);
    if constexpr (NIN & 1) sigw_single<NIN - 1>(cl, xb, lb, s);
}
// C entries from two prefetched chunk entries
template <int C>
__device__ __forceinline__ void sigh_run(uint32_t gsb, const uint32_t (&ent)[2], uint32_t lb, double (&s)[4])
{
    sigh_chunk<(C < 16 ? C : 16)>(gsb, ent[0], lb, s);
    if constexpr (C > 16) sigh_chunk<C - 16>(gsb, ent[1], lb, s);
}

template <int S, int X>
__global__ void __launch_bounds__(kSigWThreads) k_sigh_sparse(SighArgs a)
{
    static_assert(S <= 32 && X <= 16, "at most two chunks of same-half entries");
    extern __shared__ double sm[];
    const MfArgs &m = a.m;
    const int n = m.n, nb = m.nb, nn = n * n;
    // XCD-aware order: the 4 sub-chunks of one alpha row run on one XCD
    const int bid = blockIdx.x, xcd = bid & 7, k = bid >> 3;
    const int ar = xcd + 8 * (k >> 2), sub = k & 3;
    if (ar >= m.na) return;
    const long long coff = (long long)blockIdx.y * a.cstride + sub * kSigWVC;
    const double *__restrict__ x = a.x + coff;
    double *__restrict__ y = a.y + coff;
    double *buf0 = sm, *buf1 = sm + kSigHMax * kSigWVC;
    double *cf0 = buf1 + kSigHMax * kSigWVC, *cf1 = cf0 + kSigGS, *dt0 = cf1 + kSigGS, *dt1 = dt0 + 128;
    int *hls = (int *)(dt1 + 128);  // slot -> beta rank of both halves
    const int tid = threadIdx.x, lane = tid & (kLanes - 1);
    const int w = __builtin_amdgcn_readfirstlane(tid / kLanes);
    // waves 0 .. kSigHWaves-1 sum; the last wave only copies the next half
    // (its outstanding copies then never delay the summing waves' own loads)
    const bool stager = w == kSigHWaves;
    const int v = lane & 15, trow = w * 4 + (lane >> 4);
    const uint32_t alpha = m.unA[ar], full = (1u << n) - 1;
    const int ka = __popc(alpha), ne = n - ka;
    const int R = 1 + ka * ne;
    constexpr uint32_t kPad = (uint32_t)(2 * kSigNN * 8) << 17;
    for (int i = tid; i < 2 * kSigHMax; i += kSigWThreads) hls[i] = a.hl[i];
    // source row r: 0 the identity row (J = ar), else the alpha single r - 1
    auto source = [&](int r, int &rA, int &qA) {
        if (r == 0) return ar;
        const int tA = r - 1, ir = tA / ne, iq = tA - ir * ne;
        uint32_t oo = alpha, ww = ~alpha & full;
        for (int i = 0; i < ir; i++) oo &= oo - 1;
        for (int i = 0; i < iq; i++) ww &= ww - 1;
        rA = __ffs(oo) - 1;
        qA = __ffs(ww) - 1;
        return (int)m.rkA[alpha ^ (1u << rA) ^ (1u << qA)];
    };
    // signed coefficient row and beta-diagonal sums of source row r (as in
    // k_sigw_sparse), by the summing waves
    auto coefficients = [&](int r, int rA, int qA, double *gs, double *dt) {
        const bool neg = r > 0 && between_parity(alpha, rA, qA);
        const double *gb = m.g + (size_t)(rA * n + qA) * nn;
        for (int e = tid; e < nn; e += kSigHWaves * kLanes) {
            double c = 0;
            if (r == 0) {
                if (e / n != e % n)
                    for (uint32_t o = alpha; o; o &= o - 1) {
                        const int xo = __ffs(o) - 1;
                        c += thr13(m.g[(size_t)(xo * n + xo) * nn + e]);
                    }
            } else {
                c = neg ? -thr13(gb[e]) : thr13(gb[e]);
            }
            gs[e] = c;
            gs[kSigNN + e] = -c;
        }
        if (tid < kSigGS - 2 * kSigNN) gs[2 * kSigNN + tid] = 0.0;
        if (tid < 128) {
            const int hi = tid >> 6, t = tid & 63;
            double d = 0;
            if (r > 0)
                for (int b = 0; b < 6; b++) {
                    const int q = b + 6 * hi;
                    if (q < n && ((t >> b) & 1)) d += neg ? -thr13(gb[q * n + q]) : thr13(gb[q * n + q]);
                }
            dt[tid] = d;
        }
    };
    // half h of source row J into dst, by the staging wave: 8 slots (1 KB)
    // per instruction, lane l -> slot 8c + l/8, vectors 2 (l mod 8) and
    // 2 (l mod 8) + 1; slot ranks come from LDS, so no load waits between
    // the copies
    auto stage = [&](int J, int h, double *dst) {
        const int cnt = h ? a.nB : a.nA;
        const double *src = x + (long long)J * nb * kLanes + (lane & 7) * 2;
        for (int c = 0; c * 8 < cnt; c++) {
            const int slot = c * 8 + (lane >> 3);
            const int beta = hls[h * kSigHMax + (slot < cnt ? slot : 0)];
            __builtin_amdgcn_global_load_lds((const void *)(src + (long long)beta * kLanes),
                                             (__attribute__((address_space(3))) void *)(dst + c * 128), 16, 0, 0);
        }
    };
    double acc[kSigHNI];
#pragma unroll
    for (int i = 0; i < kSigHNI; i++) acc[i] = 0;
    int rA = 0, qA = 0, Jcur = ar;
    if (!stager) coefficients(0, 0, 0, cf0, dt0);
    __syncthreads();  // slot ranks in LDS
    if (stager) stage(ar, 0, buf0);
    for (int ph = 0; ph < 2 * R; ph++) {
        const int r = ph >> 1, half = ph & 1;
        // the staging wave's copies of this phase's half have landed; after
        // the barrier the other buffer and coefficient row are free
        if (stager) __builtin_amdgcn_s_waitcnt(0x0f70);  // vmcnt(0)
        __syncthreads();
        if (half == 0) {
            if (stager) stage(Jcur, 1, buf1);
        } else if (r + 1 < R) {
            Jcur = source(r + 1, rA, qA);
            if (stager) stage(Jcur, 0, buf0);
            else coefficients(r + 1, rA, qA, ((r + 1) & 1) ? cf1 : cf0, ((r + 1) & 1) ? dt1 : dt0);
        }
        if (stager) continue;
        const double *gs = (r & 1) ? cf1 : cf0, *dt = (r & 1) ? dt1 : dt0;
        const double *xs = half ? buf1 : buf0;
        const uint32_t gsb = (uint32_t)(size_t)(lds_f64 *)gs;
        const uint32_t lb = (uint32_t)(size_t)(lds_f64 *)xs + v * 8;
        // entries of the next pass are loaded while this pass sums; the empty
        // asm keeps them from being hoisted out of the phase loop
        uint32_t tb = (uint32_t)(trow * kSigWNtp + v);
        asm volatile("" : "+v"(tb));
        auto load = [&](int t, uint32_t (&ent)[2], uint32_t (&meta)[3]) {
            const bool isA = t < a.TA;
            const int off = half == 0 ? 0 : (isA ? S : X);
            const int cnt = (half == 0) == isA ? S : X;  // a second chunk only when there are > 16 entries
            const uint32_t base = tb + (uint32_t)((t - trow) * kSigWNtp);
            const bool ok = t < a.T;
            ent[0] = ok ? a.ts[base + off] : kPad;
            ent[1] = ok && cnt > 16 ? a.ts[base + off + 16] : kPad;
            meta[0] = ok ? a.ts[base - v + kSigWNtp - 3] : 0xffffffffu;  // beta rank (or -1)
            meta[1] = ok ? a.ts[base - v + kSigWNtp - 2] : 0u;           // own slot byte offset
            meta[2] = ok ? a.ts[base - v + kSigWNtp - 1] : 0u;           // the string
        };
        uint32_t ent[2], meta[3], entn[2], metan[3];
        load(trow, ent, meta);
#pragma unroll
        for (int i = 0; i < kSigHNI; i++) {
            if (kSigHPass * i >= a.T) break;
            const int t = trow + kSigHPass * i;
            load(t + kSigHPass, entn, metan);
            const bool isA = __builtin_amdgcn_readfirstlane(t) < a.TA;  // the same for the wave's four targets
            const bool same = (half == 0) == isA;
            double s[4] = {0, 0, 0, 0};
            if (same) sigh_run<S>(gsb, ent, lb, s);
            else sigh_run<X>(gsb, ent, lb, s);
            acc[i] += (s[0] + s[1]) + (s[2] + s[3]);
            if (same && meta[0] != 0xffffffffu) {
                // the target's own half: beta-diagonal sums and, for the
                // identity row, the diagonal
                const double xt = lds_at(lb + meta[1]);
                acc[i] += (dt[meta[2] & 63] + dt[64 + ((meta[2] >> 6) & 63)]) * xt;
                if (r == 0) acc[i] += m.diag[(long long)ar * nb + meta[0]] * xt;
            }
            ent[0] = entn[0];
            ent[1] = entn[1];
            meta[0] = metan[0];
            meta[1] = metan[1];
            meta[2] = metan[2];
        }
    }
    if (stager) return;
#pragma unroll
    for (int i = 0; i < kSigHNI; i++) {
        const int t = trow + kSigHPass * i;
        if (t < a.T) {
            const uint32_t rank = a.ts[(size_t)t * kSigWNtp + kSigWNtp - 3];
            if (rank != 0xffffffffu) y[((long long)ar * nb + rank) * kLanes + v] = acc[i];
        }
    }
}

template __global__ void k_sigh_sparse<2, 2>(SighArgs);
template __global__ void k_sigh_sparse<6, 3>(SighArgs);
template __global__ void k_sigh_sparse<12, 4>(SighArgs);
template __global__ void k_sigh_sparse<20, 5>(SighArgs);
template __global__ void k_sigh_sparse<30, 6>(SighArgs);

template __global__ void k_sigw_sparse<1>(SigvArgs);
template __global__ void k_sigw_sparse<2>(SigvArgs);
template __global__ void k_sigw_sparse<3>(SigvArgs);
template __global__ void k_sigw_sparse<4>(SigvArgs);
template __global__ void k_sigw_sparse<5>(SigvArgs);
template __global__ void k_sigw_sparse<6>(SigvArgs);
template __global__ void k_sigw_sparse<7>(SigvArgs);
template __global__ void k_sigw_sparse<8>(SigvArgs);
template __global__ void k_sigw_sparse<9>(SigvArgs);

// Matrix-free sigma, sparse part, vector-major vectors (single vectors and
// small batches, n <= 14): the same split and entry table as k_sigw_sparse
// with one vector per workgroup, threads over output beta strings, source
// rows staged two at a time.  LDS: [coef row 0 | coef row 1 | x row 0 |
// x row 1], coef row = kSig1GS doubles, x row = kSig1MaxNb doubles.
__global__ void __launch_bounds__(kSig1Threads) k_sig1_sparse(Sig1Args a)
{
    extern __shared__ double sm[];
    const MfArgs &m = a.m;
    const int n = m.n, nb = m.nb, nn = n * n;
    const int ar = blockIdx.x;
    const double *__restrict__ x = a.x + (long long)blockIdx.y * a.vstride;
    double *__restrict__ y = a.y + (long long)blockIdx.y * a.vstride;
    char *lds = (char *)sm;
    double *gs0 = sm, *gs1 = sm + kSig1GS;
    double *xs0 = sm + 2 * kSig1GS, *xs1 = xs0 + kSig1MaxNb;
    const int tid = threadIdx.x;
    const uint32_t alpha = m.unA[ar], full = (1u << n) - 1;
    const int ka = __popc(alpha), ne = n - ka;
    const int R = 1 + ka * ne;
    double acc[kSig1BI];
#pragma unroll
    for (int i = 0; i < kSig1BI; i++) acc[i] = 0;
    for (int r0 = 0; r0 < R; r0 += 2) {
        const bool two = r0 + 1 < R;
        __syncthreads();
        for (int h = 0; h < 2; h++) {
            if (h == 1 && !two) break;
            const int r = r0 + h;
            double *gs = h ? gs1 : gs0;
            double *xs = h ? xs1 : xs0;
            for (int e = tid; e < kSig1GS; e += kSig1Threads) gs[e] = 0.0;
            __syncthreads();
            int J;
            if (r == 0) {
                J = ar;
                for (int e = tid; e < nn; e += kSig1Threads) {
                    double vj = 0;
                    if (e / n != e % n)
                        for (uint32_t o = alpha; o; o &= o - 1) {
                            const int xo = __ffs(o) - 1;
                            vj += thr13(m.g[(size_t)(xo * n + xo) * nn + e]);
                        }
                    gs[e] = vj;
                    gs[kSig1NN + e] = -vj;
                }
            } else {
                const int tA = r - 1, ir = tA / ne, iq = tA - ir * ne;
                uint32_t oo = alpha, ww = ~alpha & full;
                for (int i = 0; i < ir; i++) oo &= oo - 1;
                for (int i = 0; i < iq; i++) ww &= ww - 1;
                const int rA = __ffs(oo) - 1, qA = __ffs(ww) - 1;
                J = m.rkA[alpha ^ (1u << rA) ^ (1u << qA)];
                const bool neg = between_parity(alpha, rA, qA);
                const double *gb = m.g + (size_t)(rA * n + qA) * nn;
                for (int e = tid; e < nn; e += kSig1Threads) {
                    const double cf = neg ? -thr13(gb[e]) : thr13(gb[e]);
                    gs[e] = cf;
                    gs[kSig1NN + e] = -cf;
                }
            }
            const double *src = x + (long long)J * nb;
            for (int i = tid; i < nb; i += kSig1Threads) xs[i] = src[i];
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < kSig1BI; i++) {
            const int beta = tid + kSig1Threads * i;
            if (beta < nb) {
                const uint4 *te = (const uint4 *)(a.ts + (size_t)beta * a.ntp);
                double s0 = 0, s1 = 0;
                for (int t4 = 0; t4 < a.ntp / 4; t4++) {
                    const uint4 q = te[t4];
                    const uint32_t ee[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                        const uint32_t xb = ee[u] & 0xffffu, cb = ee[u] >> 16;
                        s0 += *(const double *)(lds + cb) * *(const double *)(lds + 2 * kSig1GS * 8 + xb);
                        if (two)
                            s1 += *(const double *)(lds + kSig1GS * 8 + cb) *
                                  *(const double *)(lds + (2 * kSig1GS + kSig1MaxNb) * 8 + xb);
                    }
                }
                acc[i] += s0 + s1;
                if (r0 == 0) acc[i] += m.diag[(long long)ar * nb + beta] * xs0[beta];
            }
        }
    }
#pragma unroll
    for (int i = 0; i < kSig1BI; i++) {
        const int beta = tid + kSig1Threads * i;
        if (beta < nb) y[(long long)ar * nb + beta] = acc[i];
    }
}

// The same with RB source rows per phase (R = 1 + k_alpha (n - k_alpha)
// rows in ceil(R / RB) phases): each thread reads its entry list once per
// phase for RB rows, and staging fills every row of a phase between one
// pair of barriers.  LDS: RB coefficient rows of kSig1GS doubles, then RB
// source rows of nb doubles.
template <int RB>
__global__ void __launch_bounds__(kSig1Threads) k_sig1r(Sig1Args a)
{
    extern __shared__ double sm[];
    const MfArgs &m = a.m;
    const int n = m.n, nb = m.nb, nn = n * n;
    const int ar = blockIdx.x;
    const double *__restrict__ x = a.x + (long long)blockIdx.y * a.vstride;
    double *__restrict__ y = a.y + (long long)blockIdx.y * a.vstride;
    double *gsb = sm, *xsb = sm + RB * kSig1GS;
    const char *gsc = (const char *)gsb, *xsc = (const char *)xsb;
    const int tid = threadIdx.x;
    const uint32_t alpha = m.unA[ar], full = (1u << n) - 1;
    const int ka = __popc(alpha), ne = n - ka;
    const int R = 1 + ka * ne;
    double acc[kSig1BI];
#pragma unroll
    for (int i = 0; i < kSig1BI; i++) acc[i] = 0;
    const int blo = a.half ? ar : 0;  // the first target this block stores
    for (int r0 = 0; r0 < R; r0 += RB) {
        const int nr = min(RB, R - r0);
        __syncthreads();  // the previous phase's rows are no longer read
        for (int h = 0; h < nr; h++) {
            const int r = r0 + h;
            double *gs = gsb + h * kSig1GS;
            double *xs = xsb + (size_t)h * nb;
            int J = ar;
            bool neg = false;
            const double *gb = nullptr;
            if (r > 0) {
                const int tA = r - 1, ir = tA / ne, iq = tA - ir * ne;
                uint32_t oo = alpha, ww = ~alpha & full;
                for (int i = 0; i < ir; i++) oo &= oo - 1;
                for (int i = 0; i < iq; i++) ww &= ww - 1;
                const int rA = __ffs(oo) - 1, qA = __ffs(ww) - 1;
                J = m.rkA[alpha ^ (1u << rA) ^ (1u << qA)];
                neg = between_parity(alpha, rA, qA);
                gb = m.g + (size_t)(rA * n + qA) * nn;
            }
            for (int e = tid; e < kSig1GS; e += kSig1Threads) {
                const int ee = e < kSig1NN ? e : e - kSig1NN;
                double v = 0;
                if (e < 2 * kSig1NN && ee < nn) {
                    double cf = 0;
                    if (r == 0) {
                        if (ee / n != ee % n)
                            for (uint32_t o = alpha; o; o &= o - 1) {
                                const int xo = __ffs(o) - 1;
                                cf += thr13(m.g[(size_t)(xo * n + xo) * nn + ee]);
                            }
                    } else {
                        cf = neg ? -thr13(gb[ee]) : thr13(gb[ee]);
                    }
                    v = e < kSig1NN ? cf : -cf;
                }
                gs[e] = v;
            }
            const double *src = x + (long long)J * nb;
            for (int i = tid; i < nb; i += kSig1Threads) xs[i] = src[i];
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < kSig1BI; i++) {
            const int beta = tid + kSig1Threads * i;
            if (beta < nb && beta >= blo) {
                const uint4 *te = (const uint4 *)(a.ts + (size_t)beta * a.ntp);
                double sr[RB];
#pragma unroll
                for (int k = 0; k < RB; k++) sr[k] = 0;
                for (int t4 = 0; t4 < a.ntp / 4; t4++) {
                    const uint4 q = te[t4];
                    const uint32_t ee[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                        const uint32_t xb = ee[u] & 0xffffu, cb = ee[u] >> 16;
#pragma unroll
                        for (int k = 0; k < RB; k++)
                            if (k < nr)
                                sr[k] += *(const double *)(gsc + (size_t)k * kSig1GS * 8 + cb) *
                                         *(const double *)(xsc + (size_t)k * nb * 8 + xb);
                    }
                }
                double t = 0;
#pragma unroll
                for (int k = 0; k < RB; k++) t += sr[k];
                acc[i] += t;
                if (r0 == 0) acc[i] += m.diag[(long long)ar * nb + beta] * xsb[beta];
            }
        }
    }
#pragma unroll
    for (int i = 0; i < kSig1BI; i++) {
        const int beta = tid + kSig1Threads * i;
        if (beta < nb && beta >= blo) y[(long long)ar * nb + beta] = acc[i];
    }
}

template __global__ void k_sig1r<2>(Sig1Args);
template __global__ void k_sig1r<4>(Sig1Args);
template __global__ void k_sig1r<6>(Sig1Args);
template __global__ void k_sig1r<8>(Sig1Args);

// part[v * G + block] = block's share of <x_v|y_v>, vectors at stride vstride.
__global__ void __launch_bounds__(kBlock)
    k_dot_rows(const double *__restrict__ x, const double *__restrict__ y, long long N, long long vstride,
               double *__restrict__ part)
{
    __shared__ double red[kBlock / 64];
    const long long o = (long long)blockIdx.y * vstride;
    double s = 0;
    for (long long i = blockIdx.x * (long long)kBlock + threadIdx.x; i < N; i += (long long)gridDim.x * kBlock)
        s += x[o + i] * y[o + i];
    const double t = block_sum(s, red);
    if (threadIdx.x == 0) part[(long long)blockIdx.y * gridDim.x + blockIdx.x] = t;
}

// Tile (a block bx, b block by) with by <= bx: the source rows b of the
// tile's columns a are read coalesced into LDS and written transposed.  A
// block reads only entries b < a of other tiles (and its own diagonal tile
// before the barrier), and writes only entries b < a of its own.
__global__ void __launch_bounds__(kBlock) k_mirror_rows(double *__restrict__ y, int nb, long long vstride)
{
    __shared__ double t[32][33];
    const int bx = blockIdx.x, by = blockIdx.y;
    if (by > bx) return;
    double *yv = y + (long long)blockIdx.z * vstride;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int r = ty; r < 32; r += kBlock / 32) {
        const int b = by * 32 + r, a = bx * 32 + tx;
        if (b < nb && a < nb) t[r][tx] = yv[(long long)b * nb + a];
    }
    __syncthreads();
    for (int r = ty; r < 32; r += kBlock / 32) {
        const int a = bx * 32 + r, b = by * 32 + tx;
        if (a < nb && b < a) yv[(long long)a * nb + b] = t[tx][r];
    }
}

// One wave per (a, b) pair at a time, four pairs in flight per wave; each
// pair is one 512-byte line of 64 vectors read and one written.
__global__ void __launch_bounds__(kBlock) k_mirror_lanes(double *__restrict__ y, int nb, long long cstride)
{
    constexpr int W = kBlock / kLanes, U = 4;
    const int a = blockIdx.x, lane = threadIdx.x & (kLanes - 1), w = threadIdx.x / kLanes;
    double *yc = y + (long long)blockIdx.y * cstride;
    for (int b0 = w; b0 < a; b0 += W * U) {
        double v[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int b = b0 + u * W;
            v[u] = b < a ? yc[((long long)b * nb + a) * kLanes + lane] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int b = b0 + u * W;
            if (b < a) yc[((long long)a * nb + b) * kLanes + lane] = v[u];
        }
    }
}

// Per-lane dot products of lane-layout chunks: part[(c * G + block) * 64 + l].
__global__ void __launch_bounds__(kBlock)
    k_dot_lanes(const double *__restrict__ x, const double *__restrict__ y, long long N, double *__restrict__ part)
{
    constexpr int G = kBlock / kLanes;
    __shared__ double red[G][kLanes];
    const int l = threadIdx.x & (kLanes - 1), g = threadIdx.x / kLanes, c = blockIdx.y;
    const long long base = (long long)c * N * kLanes + l;
    double s = 0;
    for (long long i = blockIdx.x * (long long)G + g; i < N; i += (long long)gridDim.x * G)
        s += x[base + i * kLanes] * y[base + i * kLanes];
    red[g][l] = s;
    __syncthreads();
    if (g == 0) {
        double t = 0;
#pragma unroll
        for (int w = 0; w < G; w++) t += red[w][l];
        part[((long long)c * gridDim.x + blockIdx.x) * kLanes + l] = t;
    }
}

// ------------------------------------------------------------- Hessian
// Forward tangent step for operator k.  psi, lambda are rotated out of place
// (every grid.y slice recomputes them in registers; slice 0 writes them),
// each tangent accumulator B_mu (mu < nactive) in place; B_{injectMu} gets
// scale_k * G_k psi'; per block: X[mu] = <lambda'|G_k|B'_mu> and
// D = <lambda'|G_k G_k|psi'> = -sum_pairs (lambda'_b psi'_b + lambda'_b' psi'_b').
template <int MU>
__global__ void __launch_bounds__(kBlock)
    k_tangent(const double *__restrict__ psiIn, double *__restrict__ psiOut, const double *__restrict__ lamIn,
              double *__restrict__ lamOut, double *__restrict__ B, long long bstride, int nactive, long long N,
              DevSector sec, const DevOp *__restrict__ ops, int k, const double2 *__restrict__ cs, int injectMu,
              double scalek, double *__restrict__ xpart, int G, double *__restrict__ dpart)
{
    __shared__ double red[kBlock / 64];
    const DevOp o = ops[k];
    const double C = cs[k].x, S = cs[k].y;
    const int mu0 = blockIdx.y * MU;
    const bool lead = blockIdx.y == 0;
    double acc[MU];
#pragma unroll
    for (int t = 0; t < MU; t++) acc[t] = 0;
    double dacc = 0;
    for (long long i = blockIdx.x * (long long)kBlock + threadIdx.x; i < N; i += (long long)G * kBlock) {
        uint32_t b2;
        const uint32_t b = sec.states[i];
        const int s = pair_sign(o, b, b2);
        long long j = -1;
        if (s != 0) {
            if (b2 < b) continue;
            j = sec.rank(b2);
        }
        if (j < 0) {
            if (lead) {
                psiOut[i] = psiIn[i];
                lamOut[i] = lamIn[i];
            }
            continue;
        }
        const double pi = psiIn[i], pj = psiIn[j], li = lamIn[i], lj = lamIn[j];
        const double pi2 = s * pj * S + pi * C, pj2 = -s * pi * S + pj * C;
        const double li2 = s * lj * S + li * C, lj2 = -s * li * S + lj * C;
        if (lead) {
            psiOut[i] = pi2;
            psiOut[j] = pj2;
            lamOut[i] = li2;
            lamOut[j] = lj2;
            dacc -= li2 * pi2 + lj2 * pj2;
        }
#pragma unroll
        for (int t = 0; t < MU; t++) {
            const int mu = mu0 + t;
            if (mu >= nactive) break;
            double *Bm = B + mu * bstride;
            const double bi = Bm[i], bj = Bm[j];
            double bi2 = s * bj * S + bi * C, bj2 = -s * bi * S + bj * C;
            if (mu == injectMu) {
                bi2 += scalek * s * pj2;
                bj2 -= scalek * s * pi2;
            }
            Bm[i] = bi2;
            Bm[j] = bj2;
            acc[t] += s * (li2 * bj2 - lj2 * bi2);
        }
    }
#pragma unroll
    for (int t = 0; t < MU; t++) {
        const double tot = block_sum(acc[t], red);
        if (threadIdx.x == 0 && mu0 + t < nactive) xpart[(long long)(mu0 + t) * G + blockIdx.x] = tot;
    }
    if (lead) {
        const double tot = block_sum(dacc, red);
        if (threadIdx.x == 0) dpart[blockIdx.x] = tot;
    }
}

template __global__ void k_tangent<4>(const double *, double *, const double *, double *, double *, long long, int,
                                      long long, DevSector, const DevOp *, int, const double2 *, int, double,
                                      double *, int, double *);

__global__ void __launch_bounds__(kBlock)
    k_gram(const double *__restrict__ A, const double *__restrict__ B, long long stride, int P, long long N,
           int chunks, double *__restrict__ part)
{
    __shared__ double sa[16][65];
    __shared__ double sb[16][65];
    const int tm = blockIdx.x * 16, tn = blockIdx.y * 16, chunk = blockIdx.z;
    const long long lo = N * chunk / chunks, hi = N * (chunk + 1) / chunks;
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    double acc = 0;
    for (long long base = lo; base < hi; base += 64) {
        for (int t = threadIdx.x; t < 16 * 64; t += kBlock) {
            const int r = t >> 6, c = t & 63;
            const long long i = base + c;
            sa[r][c] = (tm + r < P && i < hi) ? A[(tm + r) * stride + i] : 0.0;
            sb[r][c] = (tn + r < P && i < hi) ? B[(tn + r) * stride + i] : 0.0;
        }
        __syncthreads();
#pragma unroll 8
        for (int c = 0; c < 64; c++) acc += sa[ty][c] * sb[tx][c];
        __syncthreads();
    }
    if (tm + ty < P && tn + tx < P) part[((long long)chunk * P + tm + ty) * P + tn + tx] = acc;
}

} // namespace ae
