// Device engine of the statevector path (host side of the kernels).
//
// One HIP stream per engine (one engine per ctx).  Everything an evaluation
// touches -- psi_0, psi, H psi, the Hamiltonian, the sector maps, the
// operator table and the sin/cos table -- is resident in HBM; a call moves
// only the per-operator sin/cos (16 B per operator) in and the results out.
#include <hip/hip_runtime.h>
// rocblas_gemm_strided_batched_ex_get_solutions (solution choice of the
// same-spin DGEMMs) is part of rocBLAS's beta API
#define ROCBLAS_BETA_FEATURES_API
#define ROCBLAS_NO_DEPRECATED_WARNINGS
#include <rocblas/rocblas.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>

#include "engine.hpp"
#include "kernels.hpp"

namespace ae {

// Scratch of the no-allocation evaluation (Engine::energy_gradient_noalloc):
// one fixed device array; every GPU the library runs on holds its own copy
// until the process ends.  Host code reaches it only through the address
// hipGetSymbolAddress returns per device.
constexpr size_t kNoAllocScratchDoubles = (size_t)4 << 20;  // 32 MiB
static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

#define HCHK(expr)                                                                     \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess) {                                                        \
            err = std::string(#expr) + " failed: " + hipGetErrorString(e_);           \
            return false;                                                              \
        }                                                                              \
    } while (0)

#define LCHK()                                                                         \
    do {                                                                               \
        hipError_t e_ = hipGetLastError();                                             \
        if (e_ != hipSuccess) {                                                        \
            err = std::string("kernel launch failed: ") + hipGetErrorString(e_);      \
            return false;                                                              \
        }                                                                              \
    } while (0)

// Grow-only device buffer.
template <class T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    bool ensure(size_t cnt, std::string &err)
    {
        if (cnt <= n && p) return true;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        if (cnt == 0) cnt = 1;
        hipError_t e = hipMalloc(&p, cnt * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            err = std::string("hipMalloc of ") + std::to_string(cnt * sizeof(T)) + " bytes failed: " +
                  hipGetErrorString(e);
            return false;
        }
        n = cnt;
        return true;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    ~DBuf() { release(); }
};

struct TileSet {
    uint32_t U = 0;
    int ntiles = 0, maxRows = 0;
    DBuf<int32_t> tileStart, rows;
    DBuf<int8_t> localRow;
    DBuf<int4> desc;  // classes of at most two members: {rank0, rank1 or -1, string0, string1}
};

// Classes of one spin for a pair of disjoint bit groups (g1, g2): strings with
// the same bits outside the groups and the same electron count in each group.
struct GroupSet {
    uint32_t key1 = 0, key2 = 0, key3 = 0;  // key3: third group (three-group sets), else 0
    bool beta = false;
    int ncls = 0;
    int n2 = 0;  // three-group sets: the first n2 classes have two group-3 members
    int nm = 0;  // two-group sets: the first nm classes have two members in some group
    DBuf<int4> ranks;
    DBuf<uint32_t> info;
};

struct Seg {
    int k0, k1;
    int tileset;  // LDS tiles (>= 0); -1: generic kernel (one operator) or orbit segment
    int orbA = -1, orbB = -1;  // orbit segment: alpha / beta class sets
    long long orbits = 0;      // alpha classes x beta classes
    uint32_t ua = 0, ub = 0;   // orbit segment: its alpha / beta bits (within one spin string)
    int progOff = -1;          // orbit segment: offset of its program (two-group: of its (g, h) list)
    int two = 0;               // lane layout: two- (or three-) group orbit segment (k_orb2v, k_orb3v)
    int three = 0;             // lane layout: three-group orbit segment (k_orb3v)
    int gA = -1, gB = -1;      // two-/three-group segment: alpha / beta group sets
    uint32_t grA[3] = {0, 0, 0}, grB[3] = {0, 0, 0};
    int prog2Off = -1;
    int prog3Off = -1;         // three-group segment: offset of its run-2 words
    int split2 = 0;            // three-group segment: operators before run 2
    int grid = 1;  // lane layout: blocks of its launches
    // lane-layout forward two-group segment with reachable amplitudes only:
    // its orbit and lone lists in d_suplist (supOff < 0: every orbit) and
    // its grid; adjoint (stored states): the orbits where the state can be
    // nonzero, also at supOff (supN of them)
    int supOff = -1, supN = 0, supLOff = 0, supLN = 0, gridSup = 1;
    int diag = 0;  // a run of number operators (complex amplitudes; k_phase)
};

// the two-half lane-layout sparse sigma kernel for (same-half, cross-half)
// entry counts, or null
static void (*sigh_kernel(int S, int X))(SighArgs)
{
    if (S == 2 && X == 2) return k_sigh_sparse<2, 2>;
    if (S == 6 && X == 3) return k_sigh_sparse<6, 3>;
    if (S == 12 && X == 4) return k_sigh_sparse<12, 4>;
    if (S == 20 && X == 5) return k_sigh_sparse<20, 5>;
    if (S == 30 && X == 6) return k_sigh_sparse<30, 6>;
    return nullptr;
}

// the lane-layout sparse sigma kernel for ng groups of 4 single entries
static void (*sigw_kernel(int ng))(SigvArgs)
{
    switch (ng) {
    case 1: return k_sigw_sparse<1>;
    case 2: return k_sigw_sparse<2>;
    case 3: return k_sigw_sparse<3>;
    case 4: return k_sigw_sparse<4>;
    case 5: return k_sigw_sparse<5>;
    case 6: return k_sigw_sparse<6>;
    case 7: return k_sigw_sparse<7>;
    case 8: return k_sigw_sparse<8>;
    default: return k_sigw_sparse<9>;
    }
}

static long long binom_ll(int n, int k)
{
    if (k < 0 || k > n) return 0;
    long long r = 1;
    for (int i = 1; i <= k; i++) r = r * (n - k + i) / i;
    return r;
}

struct EngineImpl {
    hipStream_t stream = nullptr;
    int nq = 0, mode = 0, nel = 0, nup = 0, ndown = 0, n = 0;
    long long N = 0;
    int na = 0, nb = 0;
    size_t ldsMax = 65536;
    std::vector<uint32_t> states, unA, unB;
    std::vector<int32_t> rkA, rkB;
    DBuf<uint32_t> d_states, d_unA, d_unB, d_binom;
    DBuf<int32_t> d_rkA, d_rkB;

    int M = 0, P = 0;
    std::vector<DevOp> hops;
    std::vector<int> order;
    std::vector<double> scale;
    DBuf<DevOp> d_ops;
    DBuf<uint32_t> d_prog;  // orbit programs of the orbit segments (OrbArgs::prog)
    std::vector<Seg> segs;
    // segmentation of the lane-layout sweeps: adjoint (lsegs) and forward
    // (lsegsF) may fuse different numbers of orbital groups per pass
    struct LaneList {
        std::vector<Seg> segs;
        DBuf<uint32_t> lprog;             // one-group orbit programs
        DBuf<unsigned long long> prog2;   // two-group orbit programs (and run 0/1 words of three-group ones)
        DBuf<unsigned long long> prog3;   // three-group segments: run-2 words
    };
    LaneList laneA, laneF;
    // vector-major two-group segmentation (kSweepOrbit2): two-group segments
    // where operators pair up on disjoint orbital pairs, one-group orbit
    // segments and single operators otherwise; used when it has some
    // two-group segment (sweepS)
    LaneList laneS;
    bool sweepS = false;
    bool laneSAllTwo = false;  // every segment of laneS is a two-group one (Hessian tangents in lane layout)
    bool use_s() const { return sweepKind == kSweepOrbit2 && sweepS; }
    std::vector<Seg> &lsegs = laneA.segs;
    std::vector<std::unique_ptr<GroupSet>> gsets;
    int maxGroups = 1;   // adjoint: groups per spin of a lane-layout orbit segment (1 to 3; tuned)
    int maxGroupsF = 1;  // forward
    std::vector<std::unique_ptr<TileSet>> tilesets;
    int slots = 1, Ggen = 1;
    int ngeneric = 0;

    bool haveStart = false;
    // complex amplitudes: d_start, d_psi, d_phi and d_lam hold a real and an
    // imaginary plane of N each, the single-vector sequences run on the two
    // planes as two vectors with one angle table, and number operators run
    // as k_phase segments; planes() is the number of planes
    bool cplx = false;
    int planes() const { return cplx ? 2 : 1; }
    DBuf<double> d_start, d_psi, d_phi, d_lam;
    DBuf<double2> d_cs, d_csneg;
    DBuf<double> d_partial, d_epart, d_out;
    std::vector<double2> h_cs, h_csneg;

    int hkind = 0;
    long long nnz = 0;
    int W = 0;
    DBuf<int64_t> d_rowptr;
    DBuf<double> d_h, d_g, d_diag, d_FA, d_FB;  // matrix-free tables
    // lane-layout matrix-free sigma (k_sigw_sparse or k_sigh_sparse + two DGEMMs)
    bool sigv = false;  // lane-layout split sigma available (n <= 12)
    bool sig1 = false;  // vector-major split sigma available (n <= 14)
    DBuf<double> d_Vaa, d_Vbb;
    DBuf<uint32_t> d_ts, d_ts1;
    DBuf<int4> d_sigsrc;  // grouped sigma: source rows of every alpha row, ascending
    int ntp = 0;
    int ntp16 = 0;  // lane-layout table: singles per beta rounded up to 4 (rows of kSigWNtp)
    // two-half lane sigma (k_sigh_sparse, n = 2 kd): target-slot table, slot ranks
    bool sigh = false;     // available
    int sighOpt = 0;       // kOptSigmaHalves (autotuned; the one-row kernel is faster at 12 orbitals)
    bool sigg = false;     // grouped-target sigma available (12 orbitals, 6 beta electrons)
    int siggOpt = 1;       // kOptSigmaGrouped
    int gemmSol[2] = {0, 0};  // rocBLAS solution of the beta / alpha same-spin DGEMM (0: library default)
    int sighS = 0, sighX = 0, sighNA = 0, sighNB = 0, sighTA = 0, sighT = 0;
    DBuf<uint32_t> d_tsh;
    DBuf<int> d_hl;
    rocblas_handle blas = nullptr;
    DBuf<double> bl_epart;
    DBuf<double> bl_wep;  // energy rows of the first adjoint window pass
    DBuf<int32_t> d_col;
    DBuf<double> d_val;

    // batch and Hessian workspaces
    DBuf<double> b_psi, b_lam, b_partial, b_epart, b_out;
    DBuf<double2> b_cs, b_csneg;
    DBuf<double> hz_B, hz_HB, hz_psi2, hz_lam2, hz_xpart, hz_dpart, hz_gram, hz_out, hz_epart;
    DBuf<double> hz_lx, hz_ly, hz_lep;  // tangents, H B and energy partials in lane layout (Hessian, lane sigma)
    DBuf<int> hz_inj;                   // Hessian tangent steps in lane layout: tangent position per operator
    DBuf<double> hz_scl;                // and its scale
    DBuf<double2> sub_cs, sub_csneg;  // angle tables of subspace() (apart from the graph-captured batch tables)

    int bnvec = 0;  // size of the uploaded angle batch

    // lane-layout batch workspaces (see kernels.hpp)
    DBuf<double> bl_psi, bl_lam, bl_partial;
    DBuf<double2> bl_cs, bl_csneg;
    std::vector<double2> h_csl, h_cslneg;
    int lslots = 1;   // most blocks of a lane-layout adjoint launch (= stride of the bracket partial rows)
    int lslotsF = 1;  // most blocks of a lane-layout forward launch
    DBuf<int> d_lrows;  // bracket partial rows written per operator (its launch's grid)
    // stored forward states (lane layout): state s (before segment s) at
    // bl_ck + s * nch * N * 64, s = 0 .. segments
    int ckOpt = 1;        // kOptCheckpoint
    bool ckShape = false;  // forward and adjoint lists share their segments, all orbit segments
    bool ckOn = false;     // the uploaded batch uses stored states
    DBuf<double> bl_ck;
    int numCU = 1;
    // kOptReachable: lane-layout sweeps skip the amplitudes that the start
    // state and the operators so far cannot reach (exact zeros).  supDirty:
    // the lists must be recomputed (new segments or start state; supCk: the
    // stored-state mode they were made for); supGen
    // counts recomputations; the stored states are cleared once per (buffer,
    // chunks, lists) key, since the forward sweep then writes only the
    // reachable part of each
    int reachOpt = 1;
    bool supDirty = true;
    bool hessDirty = true;  // the Hessian's tile order (hessian_tiles) must be made again
    bool supCk = false;
    long long supGen = 0;
    std::vector<double> hstart;
    DBuf<int> d_suplist;
    // amplitudes per vector and step that the lane-layout sweeps move
    // (summed over segments): forward reads and writes, adjoint lambda reads
    // and writes, adjoint stored-state reads (kOptSweepAmp*)
    long long ampF = 0, ampAL = 0, ampAP = 0;
    const double *ckZeroP = nullptr;
    size_t ckZeroNch = 0;
    long long ckZeroGen = -1;

    // tunable execution choices (Engine::set_option)
    int sweepKind = kSweepOrbit;
    int laneMin = 16;  // batches of at least this many vectors use the lane layout (0: never)
    int fuseDepth = 0;  // most operators per segment (0: the kernels' own limits)
    size_t ldsDevice = 65536;
    bool useGraph = true;
    bool orbLds = false;  // orbit segments stage their alpha rows in LDS (k_orbl)
    std::vector<int8_t> ops4;  // kept to re-segment when the sweep kind changes

    // hipGraphs: the fused single-vector E+grad sequence, the batch sequence
    hipGraphExec_t egGraph = nullptr;
    bool egSym = false;  // egGraph computes half the sigma targets (exchange-symmetric calls)
    hipGraphExec_t bGraph = nullptr;
    int bGraphNvec = 0;
    bool bGraphGrad = false;
    bool bGraphLanes = false;
    // Addresses of every buffer the batch sequence reads or writes.  A
    // captured graph holds them as kernel arguments, so a buffer that
    // ensure() moved makes it stale: batch_upload drops the graphs (and bumps
    // the generation, which multi-instance graphs key on) when the list
    // changes, and batch_run re-checks it against the captured list.
    std::vector<const void *> uploadPtrs, bGraphPtrs, egGraphPtrs;
    // the same for the single-vector E+grad graph
    std::vector<const void *> eg_ptrs() const
    {
        return {d_start.p, d_psi.p, d_phi.p, d_lam.p, d_cs.p, d_csneg.p, d_partial.p, d_epart.p, d_out.p};
    }
    // The mirror image of every operator under the exchange of the alpha
    // (upper n) and beta (lower n) bits, with its sign mask mirrored and its
    // constant sign kept: an operator is its own mirror (the paired double)
    // or sits next to it, the two acting on opposite spins only (so that
    // they commute).  opsMirror when every operator has one.
    void find_mirrors()
    {
        const int M = (int)hops.size();
        const uint32_t lo = (n >= 32 ? 0u : (1u << n)) - 1u;
        auto mir = [&](uint32_t x) { return ((x & lo) << n) | (x >> n); };
        auto same = [](const DevOp &p, const DevOp &q) {
            return p.cre == q.cre && p.ann == q.ann && p.mask == q.mask && p.flip == q.flip;
        };
        auto one_spin = [&](const DevOp &p) {
            const uint32_t u = p.cre | p.ann | p.mask;
            return (u & lo) == 0 || (u >> n) == 0;
        };
        mirror.assign(M, -1);
        opsMirror = !cplx && n > 0 && 2 * n <= 32 && M > 0;
        for (int k = 0; k < M && opsMirror; k++) {
            const DevOp &o = hops[k];
            const DevOp m = {mir(o.cre), mir(o.ann), mir(o.mask), o.flip};
            if (same(m, o)) mirror[k] = k;
            else if (one_spin(o) && k + 1 < M && same(m, hops[k + 1])) mirror[k] = k + 1;
            else if (one_spin(o) && k > 0 && same(m, hops[k - 1])) mirror[k] = k - 1;
            else opsMirror = false;
        }
        for (int k = 0; k < M && opsMirror; k++)
            if (mirror[mirror[k]] != k) opsMirror = false;
    }

    // x[a][b] == x[b][a] for the start vector (index a * nb + b)
    bool symmetric_start() const
    {
        if (na != nb || (long long)hstart.size() != (long long)na * nb) return false;
        for (long long a = 0; a < na; a++)
            for (long long b = a + 1; b < nb; b++)
                if (hstart[a * nb + b] != hstart[b * nb + a]) return false;
        return true;
    }

    // whether the batch of nvec operator-angle vectors qualifies for the
    // exchange-symmetric sigma
    bool batch_symmetric(const double *angles, int nvec) const
    {
        if (exchOpt == 0 || !opsMirror || !startSym || !vSame || !sigg || !siggOpt || hkind != 3 || !sigv ||
            cplx || na != nb)
            return false;
        for (int v = 0; v < nvec; v++) {
            const double *t = angles + (size_t)v * M;
            for (int k = 0; k < M; k++)
                if (t[k] != t[mirror[k]]) return false;
        }
        return true;
    }

    // the Hessian's tangents d psi / d theta_p are exchange-symmetric when
    // the state is and each mirror pair of operators shares its parameter
    bool tangent_symmetric(const double *angles) const
    {
        if (!batch_symmetric(angles, 1)) return false;
        for (int k = 0; k < M; k++)
            if (order[k] != order[mirror[k]] || scale[k] != scale[mirror[k]]) return false;
        return true;
    }

    // the same for single-vector calls on the vector-major split sigma
    bool vec_symmetric(const double *angles) const
    {
        if (exchOpt == 0 || !opsMirror || !startSym || hkind != 3 || !sig1 || cplx || na != nb) return false;
        for (int k = 0; k < M; k++)
            if (angles[k] != angles[mirror[k]]) return false;
        return true;
    }

    std::vector<const void *> batch_ptrs() const
    {
        return {b_psi.p,      b_lam.p,   b_partial.p, b_epart.p, b_out.p,  b_cs.p,    b_csneg.p, bl_psi.p,
                bl_lam.p,     bl_partial.p, bl_epart.p, bl_cs.p,  bl_csneg.p, bl_ck.p, d_suplist.p, d_lrows.p,
                d_wtiles.p,   d_wblob.p, d_wrows.p,   d_wkp.p,     bl_wep.p,  bl_z.p};
    }

    bool use_lanes(int nvec) const { return laneMin > 0 && nvec >= laneMin; }

    // Orbit programs (OrbArgs::prog): for each orbit type and operator of an
    // orbit segment, the member pairs the operator rotates and their phases,
    // from the pair rule on representative members with the spectator bits
    // cleared.
    bool build_orbit_programs(std::string &err) { return build_orbit_programs(segs, d_prog, err); }
    bool build_orbit_programs(std::vector<Seg> &list, DBuf<uint32_t> &dprog, std::string &err)
    {
        std::vector<uint32_t> prog;
        for (Seg &sg : list) {
            sg.progOff = -1;
            if (sg.orbA < 0 || sg.two) continue;
            sg.progOff = (int)prog.size();
            orbit_program_words(sg, prog);
        }
        if (prog.empty()) prog.push_back(0);
        if (!dprog.ensure(prog.size(), err)) return false;
        HCHK(hipMemcpy(dprog.p, prog.data(), prog.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
        return true;
    }

    // The program of one orbit segment: 9 types x kOrbMaxOps words appended to prog.
    void orbit_program_words(const Seg &sg, std::vector<uint32_t> &prog) const
    {
        uint32_t bitsA[2] = {0, 0}, bitsB[2] = {0, 0};
        {
            int i = 0;
            for (uint32_t u = sg.ua; u && i < 2; u &= u - 1) bitsA[i++] = u & (~u + 1);
            i = 0;
            for (uint32_t u = sg.ub; u && i < 2; u &= u - 1) bitsB[i++] = u & (~u + 1);
        }
        const bool twoA = __builtin_popcount(sg.ua) == 2, twoB = __builtin_popcount(sg.ub) == 2;
        for (int ta = 0; ta < 3; ta++)
            for (int tb = 0; tb < 3; tb++) {
                // members in descriptor order: lower string (lower rank) first
                uint32_t al[2], bl[2];
                int Ra = 1, Rb = 1;
                if (ta == 0) {
                    al[0] = bitsA[0];
                    al[1] = bitsA[1];
                    Ra = twoA ? 2 : 1;
                } else {
                    al[0] = al[1] = (ta == 2) ? sg.ua : 0u;
                }
                if (tb == 0) {
                    bl[0] = bitsB[0];
                    bl[1] = bitsB[1];
                    Rb = twoB ? 2 : 1;
                } else {
                    bl[0] = bl[1] = (tb == 2) ? sg.ub : 0u;
                }
                uint32_t st[4];
                bool ok[4];
                for (int p = 0; p < 4; p++) {
                    const int i = p >> 1, j = p & 1;
                    ok[p] = i < Ra && j < Rb;
                    st[p] = (al[i] << n) | bl[j];
                }
                static const int PP[6] = {0, 0, 0, 1, 1, 2}, QQ[6] = {1, 2, 3, 2, 3, 3};
                for (int kk = 0; kk < kOrbMaxOps; kk++) {
                    uint32_t w = 0;
                    if (sg.k0 + kk < sg.k1) {
                        const DevOp &o = hops[sg.k0 + kk];
                        for (int r = 0; r < 6; r++) {
                            const int p = PP[r], q = QQ[r];
                            if (!ok[p] || !ok[q]) continue;
                            uint32_t b2 = 0;
                            const int sgn = host_pair_sign(o, st[p], b2);
                            if (sgn == 0 || b2 != st[q]) continue;
                            w |= 1u << r;
                            if (sgn < 0) w |= 1u << (8 + r);
                        }
                    }
                    prog.push_back(w);
                }
            }
    }

    int groupset_for(uint32_t g1, uint32_t g2, bool beta, std::string &err)
    {
        for (size_t i = 0; i < gsets.size(); i++)
            if (gsets[i]->key1 == g1 && gsets[i]->key2 == g2 && gsets[i]->key3 == 0 && gsets[i]->beta == beta)
                return (int)i;
        auto t = std::make_unique<GroupSet>();
        t->key1 = g1;
        t->key2 = g2;
        t->beta = beta;
        const std::vector<uint32_t> &un = beta ? unB : unA;
        const uint32_t U = g1 | g2;
        std::map<std::array<uint32_t, 3>, int> cls;
        std::vector<int4> r;
        std::vector<uint32_t> info;
        auto shape = [](int c) { return c == 1 ? 0u : (c == 0 ? 1u : 2u); };
        for (int rk = 0; rk < (int)un.size(); rk++) {
            const uint32_t str = un[rk];
            const int c1 = __builtin_popcount(str & g1), c2 = __builtin_popcount(str & g2);
            const std::array<uint32_t, 3> key = {str & ~U, (uint32_t)c1, (uint32_t)c2};
            auto it = cls.find(key);
            int ci;
            if (it == cls.end()) {
                ci = (int)r.size();
                cls[key] = ci;
                r.push_back(make_int4(-1, -1, -1, -1));
                info.push_back((str & ~U) | (shape(c1) << 24) | (shape(c2) << 26));
            } else {
                ci = it->second;
            }
            // member index per group: 0 = the group's lower bit set (or 0 / 2 electrons), 1 = upper bit
            const uint32_t lo1 = g1 & (~g1 + 1), lo2 = g2 & (~g2 + 1);
            const int m1 = (c1 == 1 && !(str & lo1)) ? 1 : 0, m2 = (c2 == 1 && !(str & lo2)) ? 1 : 0;
            int *rr = &r[ci].x;
            rr[m1 + 2 * m2] = rk;
        }
        // classes with two members in some group first, single-member ones
        // last (their orbits with single-member classes of the other spin are
        // lone determinants)
        {
            std::vector<int4> r2;
            std::vector<uint32_t> i2;
            for (int pass = 0; pass < 2; pass++)
                for (size_t ci = 0; ci < r.size(); ci++) {
                    const bool multi = ((info[ci] >> 24) & 3) == 0 || ((info[ci] >> 26) & 3) == 0;
                    if (multi != (pass == 0)) continue;
                    r2.push_back(r[ci]);
                    i2.push_back(info[ci]);
                    if (multi) t->nm++;
                }
            r.swap(r2);
            info.swap(i2);
        }
        t->ncls = (int)r.size();
        if (!t->ranks.ensure(r.size(), err) || !t->info.ensure(info.size(), err)) return -1;
        HCHK(hipMemcpy(t->ranks.p, r.data(), r.size() * sizeof(int4), hipMemcpyHostToDevice));
        HCHK(hipMemcpy(t->info.p, info.data(), info.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
        gsets.push_back(std::move(t));
        return (int)gsets.size() - 1;
    }

    // Class set of a three-group segment (k_orb3v): strings agreeing outside
    // g1 | g2 | g3 with the same electron count in each group; members by
    // (m1, m2, m3), m = 0 the group's lower bit (or 0 / 2 electrons), 1 the
    // upper; two int4 per class (m3 = 0 then 1, each indexed m1 + 2 m2).
    // Classes with one electron in g3 (two group-3 members) come first.
    int groupset3_for(uint32_t g1, uint32_t g2, uint32_t g3, bool beta, std::string &err)
    {
        for (size_t i = 0; i < gsets.size(); i++)
            if (gsets[i]->key1 == g1 && gsets[i]->key2 == g2 && gsets[i]->key3 == g3 && gsets[i]->beta == beta)
                return (int)i;
        auto t = std::make_unique<GroupSet>();
        t->key1 = g1;
        t->key2 = g2;
        t->key3 = g3;
        t->beta = beta;
        const std::vector<uint32_t> &un = beta ? unB : unA;
        const uint32_t U = g1 | g2 | g3;
        auto shape = [](int c) { return c == 1 ? 0u : (c == 0 ? 1u : 2u); };
        std::map<std::array<uint32_t, 4>, int> cls;
        std::vector<std::array<int, 8>> mem;
        std::vector<uint32_t> info;
        for (int rk = 0; rk < (int)un.size(); rk++) {
            const uint32_t str = un[rk];
            const int c1 = __builtin_popcount(str & g1), c2 = __builtin_popcount(str & g2),
                      c3 = __builtin_popcount(str & g3);
            const std::array<uint32_t, 4> key = {str & ~U, (uint32_t)c1, (uint32_t)c2, (uint32_t)c3};
            auto it = cls.find(key);
            int ci;
            if (it == cls.end()) {
                ci = (int)mem.size();
                cls[key] = ci;
                mem.push_back({-1, -1, -1, -1, -1, -1, -1, -1});
                info.push_back((str & ~U) | (shape(c1) << 24) | (shape(c2) << 26) | (shape(c3) << 28));
            } else {
                ci = it->second;
            }
            const uint32_t lo1 = g1 & (~g1 + 1), lo2 = g2 & (~g2 + 1), lo3 = g3 & (~g3 + 1);
            const int m1 = (c1 == 1 && !(str & lo1)) ? 1 : 0, m2 = (c2 == 1 && !(str & lo2)) ? 1 : 0,
                      m3 = (c3 == 1 && !(str & lo3)) ? 1 : 0;
            mem[ci][4 * m3 + m1 + 2 * m2] = rk;
        }
        // classes with two group-3 members first (stable within each part)
        std::vector<int> order;
        for (int pass = 0; pass < 2; pass++)
            for (int ci = 0; ci < (int)mem.size(); ci++)
                if ((((info[ci] >> 28) & 3) == 0) == (pass == 0)) order.push_back(ci);
        std::vector<int4> r;
        std::vector<uint32_t> inf;
        for (int ci : order) {
            const std::array<int, 8> &m = mem[ci];
            r.push_back(make_int4(m[0], m[1], m[2], m[3]));
            r.push_back(make_int4(m[4], m[5], m[6], m[7]));
            inf.push_back(info[ci]);
            if (((info[ci] >> 28) & 3) == 0) t->n2++;
        }
        t->ncls = (int)inf.size();
        if (!t->ranks.ensure(r.size(), err) || !t->info.ensure(inf.size(), err)) return -1;
        HCHK(hipMemcpy(t->ranks.p, r.data(), r.size() * sizeof(int4), hipMemcpyHostToDevice));
        HCHK(hipMemcpy(t->info.p, inf.data(), inf.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
        gsets.push_back(std::move(t));
        return (int)gsets.size() - 1;
    }

    // Lane-layout segmentation (S_z sector): two-group orbit segments where
    // operators pair up on disjoint spatial pairs, one-group orbit segments
    // otherwise, and the vector-major segments' LDS tiles and generic
    // operators run one operator per pass (k_opv).
    bool build_lane_segments(std::string &err)
    {
        supDirty = hessDirty = true;
        if (!build_lane_list(laneA, maxGroups, err) || !build_lane_list(laneF, maxGroupsF, err)) return false;
        sweepS = laneSAllTwo = false;
        laneS.segs.clear();
        if (mode == kSz && sweepKind != kSweepTile) {
            if (!build_lane_list(laneS, 2, err)) return false;
            sweepS = false;
            laneSAllTwo = !laneS.segs.empty();
            for (const Seg &sg : laneS.segs) {
                laneSAllTwo = laneSAllTwo && sg.two && !sg.three;
                if (sg.three) {
                    sweepS = false;
                    break;
                }
                sweepS = sweepS || sg.two;
            }
        }
        // Every launch is one grid-stride pass of exactly as many blocks as
        // the chip holds at once for that kernel (its occupancy x CUs): a
        // grid a little larger than that leaves a second, sparsely occupied
        // round of blocks that each still carry a full share of the orbits.
        // Fewer blocks only when the segment has less work than that.
        auto resident = [&](const void *fn) {
            int nb = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, kBlock, 0) != hipSuccess || nb < 1) nb = 1;
            return (long long)nb * numCU;
        };
        auto assign = [&](std::vector<Seg> &l, bool adj) {
            int mx = 1;
            for (Seg &sg : l) {
                const void *fn = sg.three ? (adj ? (const void *)k_orb3v<true> : (const void *)k_orb3v<false>)
                                 : sg.two ? (adj ? (const void *)k_orb2v<true> : (const void *)k_orb2v<false>)
                                 : sg.orbA >= 0 ? (adj ? (const void *)k_orbv<true> : (const void *)k_orbv<false>)
                                                : (adj ? (const void *)k_opv<true> : (const void *)k_opv<false>);
                // waves with work: one orbit (or one 64-determinant tile) each
                const long long work = (sg.orbA >= 0 || sg.two) ? sg.orbits : (N + kLanes - 1) / kLanes;
                const long long want = (work + kBlock / kLanes - 1) / (kBlock / kLanes);
                sg.grid = (int)std::max<long long>(1, std::min(resident(fn), want));
                mx = std::max(mx, sg.grid);
            }
            return mx;
        };
        lslots = assign(laneA.segs, true);
        lslotsF = assign(laneF.segs, false);
        ckShape = laneA.segs.size() == laneF.segs.size() && !laneA.segs.empty();
        for (size_t i = 0; ckShape && i < laneA.segs.size(); i++) {
            const Seg &a = laneA.segs[i], &f = laneF.segs[i];
            ckShape = a.k0 == f.k0 && a.k1 == f.k1 && (a.two || a.orbA >= 0) && (f.two || f.orbA >= 0);
        }
        // partial rows per operator: the grid of the adjoint launch that owns it
        std::vector<int> rows(M, 0);
        for (const Seg &sg : laneA.segs)
            for (int k = sg.k0; k < sg.k1; k++) rows[k] = sg.grid;
        if (!d_lrows.ensure(std::max(M, 1), err)) return false;
        HCHK(hipMemcpy(d_lrows.p, rows.data(), M * sizeof(int), hipMemcpyHostToDevice));
        return build_window_plan(err);
    }

    // Reachable amplitudes of the lane-layout sweeps.  S_0 = the start
    // state's nonzero determinants; a two-group forward segment s acts on
    // an orbit only if some member is in S_s (the others hold exact zeros
    // and stay zero), and S_{s+1} = S_s + every member of those orbits.  Its
    // list holds those orbits and the lone determinants in S_s.  An
    // adjoint segment (stored states, the same segments) acts only on the
    // orbits meeting S_{s+1}: elsewhere the state is zero, so the brackets
    // are, and the lambda there never meets a nonzero state again (the
    // segments after it in the adjoint order see S_s, S_{s-1}, ..., subsets
    // of S_{s+1}, which is a union of this segment's orbits).  A segment of
    // another kind makes everything reachable from there on.
    bool compute_support(bool ck, std::string &err)
    {
        if (winOn) return compute_support_win(ck, err);
        supDirty = false;
        supCk = ck;
        supGen++;
        for (Seg &sg : laneF.segs) sg.supOff = -1;
        for (Seg &sg : laneA.segs) sg.supOff = -1;
        // without lists every amplitude counts as reachable (the amplitude
        // counts are still computed)
        const bool lists_on = reachOpt && (long long)hstart.size() == N;
        std::vector<uint8_t> S(N, (uint8_t)1);
        if (lists_on)
            for (long long i = 0; i < N; i++) S[i] = hstart[i] != 0.0;
        ampF = ampAL = ampAP = 0;
        std::vector<std::vector<int4>> rk(gsets.size());
        auto ranks = [&](int g) -> const std::vector<int4> & {
            std::vector<int4> &r = rk[g];
            if (r.empty() && gsets[g]->ncls > 0) {
                r.resize(gsets[g]->ncls);
                if (hipMemcpy(r.data(), gsets[g]->ranks.p, r.size() * sizeof(int4), hipMemcpyDeviceToHost) !=
                    hipSuccess)
                    r.clear();
            }
            return r;
        };
        // the orbit index space of k_orb2v: alpha classes [0, nAm) x every
        // beta class, then [nAm, nAc) x [0, nBm); lone: [nAm, nAc) x [nBm, nBc)
        auto for_orbit = [&](const GroupSet &A, const GroupSet &B, const std::vector<int4> &ra,
                             const std::vector<int4> &rb, int t, auto &&f) {
            const int mainA = A.nm * B.ncls;
            int ia, ib;
            if (t < mainA) {
                ia = t / B.ncls;
                ib = t - ia * B.ncls;
            } else {
                const int u = t - mainA;
                ia = A.nm + u / B.nm;
                ib = u - (ia - A.nm) * B.nm;
            }
            const int a4[4] = {ra[ia].x, ra[ia].y, ra[ia].z, ra[ia].w};
            const int b4[4] = {rb[ib].x, rb[ib].y, rb[ib].z, rb[ib].w};
            for (int ma = 0; ma < 4; ma++)
                for (int mb = 0; mb < 4; mb++)
                    if (a4[ma] >= 0 && b4[mb] >= 0) f((long long)a4[ma] * nb + b4[mb]);
        };
        std::vector<int> lists;
        bool full = false;
        const bool adj = ckShape && ck;
        // adjoint lambda: every determinant but the lone ones of two-group segments
        std::vector<long long> loneA(laneA.segs.size(), 0);
        for (size_t i = 0; i < laneA.segs.size(); i++) {
            const Seg &sa = laneA.segs[i];
            if (sa.two && !sa.three) {
                const GroupSet &AA = *gsets[sa.gA], &BB = *gsets[sa.gB];
                loneA[i] = (long long)(AA.ncls - AA.nm) * (BB.ncls - BB.nm);
            }
            if (!adj || !sa.two || sa.three) {  // no mask: every acting orbit
                ampAL += N - loneA[i];
                ampAP += N - loneA[i];
            }
        }
        for (size_t si = 0; si < laneF.segs.size(); si++) {
            Seg &sg = laneF.segs[si];
            if (!full && (!sg.two || sg.three)) {
                full = true;
                std::fill(S.begin(), S.end(), (uint8_t)1);
            }
            if (full) {
                ampF += N;
                if (adj && si < laneA.segs.size() && laneA.segs[si].two && !laneA.segs[si].three) {
                    ampAL += N - loneA[si];
                    ampAP += N - loneA[si];
                }
                continue;
            }
            const GroupSet &A = *gsets[sg.gA], &B = *gsets[sg.gB];
            const std::vector<int4> &ra = ranks(sg.gA), &rb = ranks(sg.gB);
            if (ra.empty() || rb.empty()) {
                err = "reading class ranks failed";
                return false;
            }
            const int total = A.nm * B.ncls + (A.ncls - A.nm) * B.nm;
            std::vector<int> act;
            for (int t = 0; t < total; t++) {
                bool any = false;
                for_orbit(A, B, ra, rb, t, [&](long long d) { any = any || S[d]; });
                if (any) act.push_back(t);
            }
            const int nBl = B.ncls - B.nm, nl = (A.ncls - A.nm) * nBl;
            std::vector<int> lone;
            for (int t = 0; t < nl; t++) {
                const int i = t / nBl, ia = A.nm + i, ib = B.nm + (t - i * nBl);
                if (S[(long long)ra[ia].x * nb + rb[ib].x]) lone.push_back(t);
            }
            for (int t : act) for_orbit(A, B, ra, rb, t, [&](long long d) {
                    S[d] = 1;
                    ampF++;
                });
            ampF += (long long)lone.size();
            if ((int)act.size() < total || (int)lone.size() < nl) {
                sg.supOff = (int)lists.size();
                sg.supN = (int)act.size();
                lists.insert(lists.end(), act.begin(), act.end());
                sg.supLOff = (int)lists.size();
                sg.supLN = (int)lone.size();
                lists.insert(lists.end(), lone.begin(), lone.end());
                const int wpb = kBlock / kLanes;
                const long long want = (std::max<long long>(sg.supN, (sg.supLN + 3) / 4) + wpb - 1) / wpb;
                sg.gridSup = (int)std::max<long long>(1, std::min<long long>(sg.grid, want));
            }
            // the adjoint segment reading the state this one writes
            if (adj && si < laneA.segs.size() && laneA.segs[si].two && !laneA.segs[si].three) {
                Seg &sa = laneA.segs[si];
                const GroupSet &AA = *gsets[sa.gA], &BB = *gsets[sa.gB];
                const std::vector<int4> &raa = ranks(sa.gA), &rbb = ranks(sa.gB);
                const int tot = AA.nm * BB.ncls + (AA.ncls - AA.nm) * BB.nm;
                std::vector<int> liveOrb;
                for (int t = 0; t < tot; t++) {
                    bool any = false;
                    long long cnt = 0;
                    for_orbit(AA, BB, raa, rbb, t, [&](long long d) {
                        any = any || S[d];
                        cnt++;
                    });
                    if (any) {
                        liveOrb.push_back(t);
                        ampAL += cnt;
                        ampAP += cnt;
                    }
                }
                if ((int)liveOrb.size() < tot && lists_on) {
                    sa.supOff = (int)lists.size();
                    sa.supN = (int)liveOrb.size();
                    lists.insert(lists.end(), liveOrb.begin(), liveOrb.end());
                }
            }
        }
        if (!lists.empty()) {
            if (!d_suplist.ensure(lists.size(), err)) return false;
            HCHK(hipMemcpy(d_suplist.p, lists.data(), lists.size() * sizeof(int), hipMemcpyHostToDevice));
        }
        return true;
    }

    // ---------------------------------------- window passes (kOptWindow)
    // A pass applies every orbit segment ("phase": consecutive operators on
    // at most two alpha and two beta spin-orbitals) that lies inside a window
    // of w consecutive spatial orbitals and commutes with every phase it
    // overtakes (disjoint spin-orbitals; even fermionic operators on disjoint
    // modes commute), to LDS-staged tiles (k_win).  The passes are chosen
    // greedily with one step of look-ahead for the fewest passes.
    struct WinPass {
        int w0 = 0;
        int nph = 0, kp0 = 0, nkp = 0;
        int nAcls = 0, nBcls = 0;
        int blobOff = 0, nblob = 0;  // its table blob in d_wblob (WinArgs)
        int oSubIdx = 0, oSub = 0, oProg = 0, oMask = 0, oPat = 0, oCOff = 0, oCls = 0, oCStart = 0, oProgT = 0;
        int rowsOff = 0, nrowsA = 0;  // its class member lists in d_wrows
        std::vector<int> cntA, cntB, startA, startB;  // class sizes and first members (per spin)
        std::vector<int> incA, incB;                  // inner electron counts of the classes
        std::vector<uint32_t> outA, outB;             // their outer bits
        int maxMembers = 0, maxSide = 0;
        std::vector<int> allTiles;   // every tile, largest first
        std::vector<int> tileSize;   // members per tile (indexed by tile id)
        int tileOff = 0, ntiles = 0; // the tiles a sweep processes (d_wtiles)
        int gridF = 1, gridA = 1;    // blocks (multiples of 64 / V)
        // vector-major register tiles (k_winv): every tile, grouped by class,
        // in tasks of at most 64 (d_wvtasks), each with its record block in
        // d_wvrec (words from vRecOff)
        int vRecOff = 0, vTaskOff = 0, vTasks = 0;
        int gridVF = 1, gridVA = 1;
        // staged alpha rows (k_winvs): beta-side tasks and their record
        // blocks (d_wvsrec, words from vsRecOff), waves per block, grids, LDS
        int vsRecOff = 0, vsTaskOff = 0, vsTasks = 0, vsWavesF = 1, vsWavesA = 1, gridSF = 1, gridSA = 1;
        size_t vsLdsF = 0, vsLdsA = 0;
        int allOff = 0;  // every tile in d_wall (Hessian tangent passes, k_wint)
        int nReach = 0;  // ... the first nReach of them reachable from the start state
        int gridT = 1;
    };
    // kOptWindow: window width in spatial orbitals (0: off; -1, the default:
    // four-orbital windows when their tiles fit in registers, else off)
    int winW = -1;
    int winWaves = 4;    // kOptWindowWaves: waves per block of the window kernels (4 or 8)
    bool winOn = false;  // a plan exists: lane-layout sweeps run as window passes
    // the plan's windows are four orbitals wide and every phase is a TUPS
    // unit (the same adjacent pair in both spins): tiles in registers (k_winr)
    bool winReg = false;
    // ... and every phase is a TUPS unit of five operators whose programs
    // rotate exactly that unit's member pairs (straight-line phases, k_winr TU)
    bool winTu = false;
    std::vector<WinPass> wpass;
    DBuf<uint32_t> d_wblob, d_wrows;
    std::vector<uint32_t> hrows;  // host copy of d_wrows
    DBuf<int> d_wkp, d_wtiles, d_wall;
    DBuf<uint32_t> d_wvrec, d_wvsrec;
    DBuf<int4> d_wvstasks;
    bool winVsOK = false;  // every pass fits k_winvs's LDS
    bool winVs = true;     // kOptWindowStaged
    // Exchange symmetry of the alpha and beta strings (kOptSigmaExchange).
    // With the same strings per spin, x[a][b] = x[b][a] for every state of
    // the batch when the start vector has it and the unitary commutes with
    // the exchange; then the beta-beta part of the same-spin sigma,
    // sum_j V[b][j] x[a][j], equals (V x)[b][a], the alpha-alpha DGEMM's
    // result read transposed.
    int exchOpt = -1;            // -1 automatic, 0 off
    std::vector<int> mirror;     // per operator: itself, or the neighbour that is its mirror image
    bool opsMirror = false;      // every operator has one (set_operators)
    bool startSym = false;       // the start vector is exchange-symmetric (set_start)
    bool vSame = false;          // V_aa == V_bb bit for bit (sigma set-up)
    bool bSym = false;           // the uploaded batch qualifies
    bool bSymV = false;          // the uploaded vector-major batch qualifies (vec_symmetric per vector)
    DBuf<double> bl_z;           // alpha-alpha DGEMM result of the batch (bSym)
    DBuf<double> hz_z;           // the same for the Hessian's tangents
    DBuf<int4> d_wvtasks;
    // kOptWindowVector: window passes for the vector-major sweeps (-1, the
    // default: with register tiles from 2^18 determinants; 0 off; 1 with
    // register tiles at any size)
    int winVec = -1;
    bool winFuseE = true;  // kOptWindowEnergy
    bool winHess = true;   // kOptWindowHessian
    bool use_wv() const
    {
        return winOn && winReg && !cplx && (winVec > 0 || (winVec < 0 && N >= (1LL << 18)));
    }
    int wslots = 1;  // bracket partial rows of the window adjoint passes

    int nfwd() const { return winOn ? (int)wpass.size() : (int)laneF.segs.size(); }

    template <class T>
    bool upload_vec(DBuf<T> &d, const std::vector<T> &h, std::string &err)
    {
        if (!d.ensure(std::max<size_t>(h.size(), 1), err)) return false;
        if (!h.empty()) HCHK(hipMemcpy(d.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
        return true;
    }

    const void *winr_fn(bool br, bool bc) const
    {
        if (winTu)
            return br ? (const void *)k_winr<true, false, 2, true>
                      : (bc ? (const void *)k_winr<false, true, 4, true> : (const void *)k_winr<false, false, 4, true>);
        return br ? (const void *)k_winr<true, false, 2, false>
                  : (bc ? (const void *)k_winr<false, true, 4, false> : (const void *)k_winr<false, false, 4, false>);
    }
    int win_waves(bool br) const { return winReg ? (br ? 2 : 4) : winWaves; }
    const void *winvs_fn(bool br) const
    {
        return winTu ? (br ? (const void *)k_winvs<true, true> : (const void *)k_winvs<false, true>)
                     : (br ? (const void *)k_winvs<true, false> : (const void *)k_winvs<false, false>);
    }
    bool use_vs() const { return winVs && winVsOK; }
    const void *winv_fn(bool br) const
    {
        return winTu ? (br ? (const void *)k_winv<true, true> : (const void *)k_winv<false, true>)
                     : (br ? (const void *)k_winv<true, false> : (const void *)k_winv<false, false>);
    }

    // Window passes of the vector-major sweeps (k_winv): in place, nvec
    // vectors N apart, angle tables M apart per vector.
    bool sweep_forward_wv(double *v, int nvec, const double2 *cs, std::string &err)
    {
        const int mk = mark_begin(1);
        for (const WinPass &wp : wpass) {
            const WinArgs a = winargs(wp);
            if (use_vs()) {
                const dim3 grid((unsigned)wp.gridSF, nvec), blk(wp.vsWavesF * kLanes);
                if (winTu)
                    hipLaunchKernelGGL((k_winvs<false, true>), grid, blk, wp.vsLdsF, stream, a,
                                       d_wvstasks.p + wp.vsTaskOff, wp.vsTasks, d_wvsrec.p + wp.vsRecOff, nb, v, nullptr,
                                       N, cs, M, nullptr, 0);
                else
                    hipLaunchKernelGGL((k_winvs<false, false>), grid, blk, wp.vsLdsF, stream, a,
                                       d_wvstasks.p + wp.vsTaskOff, wp.vsTasks, d_wvsrec.p + wp.vsRecOff, nb, v, nullptr,
                                       N, cs, M, nullptr, 0);
                LCHK();
                continue;
            }
            const dim3 grid((unsigned)wp.gridVF, nvec);
            if (winTu)
                hipLaunchKernelGGL((k_winv<false, true>), grid, dim3(4 * kLanes), 0, stream, a,
                                   d_wvtasks.p + wp.vTaskOff, wp.vTasks, d_wvrec.p + wp.vRecOff, v, nullptr, N, cs,
                                   M, nullptr, 0);
            else
                hipLaunchKernelGGL((k_winv<false, false>), grid, dim3(4 * kLanes), 0, stream, a,
                                   d_wvtasks.p + wp.vTaskOff, wp.vTasks, d_wvrec.p + wp.vRecOff, v, nullptr, N, cs,
                                   M, nullptr, 0);
            LCHK();
        }
        mark_end(mk);
        return true;
    }
    // partial: nvec x M rows of `slots` (zeroed by the caller); block b of a
    // pass writes row b of the pass's operators
    bool sweep_reverse_wv(double *phi, double *lam, int nvec, const double2 *csneg, double *partial,
                          std::string &err)
    {
        const int mk = mark_begin(2);
        for (int p = (int)wpass.size() - 1; p >= 0; p--) {
            const WinPass &wp = wpass[p];
            const WinArgs a = winargs(wp);
            if (use_vs()) {
                const dim3 grid((unsigned)wp.gridSA, nvec), blk(wp.vsWavesA * kLanes);
                if (winTu)
                    hipLaunchKernelGGL((k_winvs<true, true>), grid, blk, wp.vsLdsA, stream, a,
                                       d_wvstasks.p + wp.vsTaskOff, wp.vsTasks, d_wvsrec.p + wp.vsRecOff, nb, phi, lam,
                                       N, csneg, M, partial, slots);
                else
                    hipLaunchKernelGGL((k_winvs<true, false>), grid, blk, wp.vsLdsA, stream, a,
                                       d_wvstasks.p + wp.vsTaskOff, wp.vsTasks, d_wvsrec.p + wp.vsRecOff, nb, phi, lam,
                                       N, csneg, M, partial, slots);
                LCHK();
                continue;
            }
            const size_t lds = (size_t)4 * wp.nkp * kLanes * sizeof(double);
            const dim3 grid((unsigned)wp.gridVA, nvec);
            if (winTu)
                hipLaunchKernelGGL((k_winv<true, true>), grid, dim3(4 * kLanes), lds, stream, a,
                                   d_wvtasks.p + wp.vTaskOff, wp.vTasks, d_wvrec.p + wp.vRecOff, phi, lam, N, csneg,
                                   M, partial, slots);
            else
                hipLaunchKernelGGL((k_winv<true, false>), grid, dim3(4 * kLanes), lds, stream, a,
                                   d_wvtasks.p + wp.vTaskOff, wp.vTasks, d_wvrec.p + wp.vRecOff, phi, lam, N, csneg,
                                   M, partial, slots);
            LCHK();
        }
        mark_end(mk);
        return true;
    }
    static const void *win_fn(bool br, int nw, bool bc)
    {
        if (nw == 8) return br ? (const void *)k_win<true, false, 8>
                               : (bc ? (const void *)k_win<false, true, 8> : (const void *)k_win<false, false, 8>);
        return br ? (const void *)k_win<true, false, 4>
                  : (bc ? (const void *)k_win<false, true, 4> : (const void *)k_win<false, false, 4>);
    }

    bool build_window_plan(std::string &err)
    {
        winOn = false;
        winReg = winTu = false;
        wpass.clear();
        if (winW == 0 || mode != kSz || cplx || M == 0) return true;
        const int w = std::min(std::min(winW < 0 ? 4 : winW, kWinMaxW), n);
        const uint32_t amask = (1u << n) - 1;
        auto bitsA = [&](int kk) { return ((hops[kk].cre | hops[kk].ann) >> n) & amask; };
        auto bitsB = [&](int kk) { return (hops[kk].cre | hops[kk].ann) & amask; };
        struct Ph {
            int k0, k1;
            uint32_t ua, ub;
        };
        std::vector<Ph> phs;
        for (int k = 0; k < M;) {
            uint32_t ua = bitsA(k), ub = bitsB(k);
            if (hops[k].cre == hops[k].ann || __builtin_popcount(ua) > 2 || __builtin_popcount(ub) > 2) return true;
            auto span = [](uint32_t sp) { return (31 - __builtin_clz(sp)) - __builtin_ctz(sp) + 1; };
            if (span(ua | ub) > w) return true;  // an operator wider than the window
            int k1 = k + 1;
            while (k1 < M && k1 - k < kOrbMaxOps && hops[k1].cre != hops[k1].ann) {
                const uint32_t a2 = ua | bitsA(k1), b2 = ub | bitsB(k1);
                if (__builtin_popcount(a2) > 2 || __builtin_popcount(b2) > 2 || span(a2 | b2) > w) break;
                ua = a2;
                ub = b2;
                k1++;
            }
            phs.push_back({k, k1, ua, ub});
            k = k1;
        }
        const int nw = n - w + 1;
        const size_t P = phs.size();
        // the phases window w0 can take now: inside it, not overtaking an
        // unapplied phase on a shared spin-orbital, within the pass limits
        auto pick = [&](const std::vector<char> &done, int w0, std::vector<int> &list) {
            const uint32_t W = ((1u << w) - 1) << w0;
            uint32_t tA = 0, tB = 0;
            int ops = 0;
            list.clear();
            for (size_t i = 0; i < P; i++) {
                if (done[i]) continue;
                const Ph &p = phs[i];
                const int nk = p.k1 - p.k0;
                if (!((p.ua | p.ub) & ~W) && !(p.ua & tA) && !(p.ub & tB) && (int)list.size() < kWinMaxPhases &&
                    ops + nk <= kWinMaxOps) {
                    list.push_back((int)i);
                    ops += nk;
                } else {
                    tA |= p.ua;
                    tB |= p.ub;
                }
            }
        };
        auto greedy = [&](std::vector<char> done, size_t left) {
            int passes = 0;
            std::vector<int> l, best;
            while (left) {
                best.clear();
                for (int w0 = 0; w0 < nw; w0++) {
                    pick(done, w0, l);
                    if (l.size() > best.size()) best = l;
                }
                if (best.empty()) return 1 << 20;
                for (int i : best) done[i] = 1;
                left -= best.size();
                passes++;
            }
            return passes;
        };
        std::vector<char> done(P, 0);
        size_t left = P;
        std::vector<std::pair<int, std::vector<int>>> plan;
        while (left) {
            int bestCost = 1 << 30, bestW = -1;
            std::vector<int> bestL, l;
            for (int w0 = 0; w0 < nw; w0++) {
                pick(done, w0, l);
                if (l.empty()) continue;
                std::vector<char> d2 = done;
                for (int i : l) d2[i] = 1;
                const int cost = 1 + greedy(d2, left - l.size());
                if (cost < bestCost || (cost == bestCost && l.size() > bestL.size())) {
                    bestCost = cost;
                    bestW = w0;
                    bestL = l;
                }
            }
            if (bestW < 0) return true;
            for (int i : bestL) done[i] = 1;
            left -= bestL.size();
            plan.push_back({bestW, bestL});
        }
        // Fewer passes, if there are: breadth first over the sets of applied
        // phases, a level's successors being what each window position takes
        // from each of its sets; each level keeps the sets that no other kept
        // set contains, at most 256, the fullest first (more applied phases
        // never leave a window less to take).  Stops at one pass short of the
        // plan above; its path of window positions is replayed.
        {
            const size_t NWd = (P + 63) / 64;
            using Bits = std::vector<uint64_t>;
            struct Node {
                Bits d;
                int parent, w0;
            };
            std::vector<std::vector<Node>> lv(1);
            lv[0].push_back({Bits(NWd, 0), -1, -1});
            std::vector<char> dc(P);
            std::vector<int> l;
            int found = -1;  // index of the completing node in the last level
            for (size_t depth = 1; depth < plan.size() && found < 0; depth++) {
                std::map<Bits, std::pair<int, int>> next;  // set -> (parent, window)
                const std::vector<Node> &cur = lv.back();
                for (size_t ni = 0; ni < cur.size() && found < 0; ni++) {
                    for (size_t i = 0; i < P; i++) dc[i] = (char)((cur[ni].d[i >> 6] >> (i & 63)) & 1);
                    for (int w0 = 0; w0 < nw; w0++) {
                        pick(dc, w0, l);
                        if (l.empty()) continue;
                        Bits d2 = cur[ni].d;
                        for (int i : l) d2[(size_t)i >> 6] |= 1ull << (i & 63);
                        next.emplace(d2, std::make_pair((int)ni, w0));
                    }
                }
                std::vector<Node> kept;
                std::vector<std::pair<int, const Bits *>> order;
                for (const auto &kv : next) {
                    int c = 0;
                    for (uint64_t x : kv.first) c += __builtin_popcountll(x);
                    order.push_back({c, &kv.first});
                }
                std::stable_sort(order.begin(), order.end(), [](const auto &x, const auto &y) { return x.first > y.first; });
                for (const auto &o : order) {
                    if ((size_t)o.first == P) {  // every phase applied
                        const auto &pw = next[*o.second];
                        kept.clear();
                        kept.push_back({*o.second, pw.first, pw.second});
                        found = 0;
                        break;
                    }
                    bool dominated = false;
                    for (const Node &k : kept) {
                        bool sub = true;
                        for (size_t q = 0; q < NWd && sub; q++) sub = (o.second->at(q) & ~k.d[q]) == 0;
                        if (sub) {
                            dominated = true;
                            break;
                        }
                    }
                    if (dominated) continue;
                    const auto &pw = next[*o.second];
                    kept.push_back({*o.second, pw.first, pw.second});
                    if (kept.size() >= 256) break;
                }
                if (kept.empty()) break;
                lv.push_back(std::move(kept));
            }
            if (found >= 0) {
                std::vector<int> ws;
                for (int li = (int)lv.size() - 1, ni = found; li > 0; li--) {
                    ws.push_back(lv[li][ni].w0);
                    ni = lv[li][ni].parent;
                }
                std::reverse(ws.begin(), ws.end());
                std::vector<std::pair<int, std::vector<int>>> plan2;
                std::fill(dc.begin(), dc.end(), (char)0);
                for (int w0 : ws) {
                    pick(dc, w0, l);
                    for (int i : l) dc[i] = 1;
                    plan2.push_back({w0, l});
                }
                if (std::all_of(dc.begin(), dc.end(), [](char c) { return c != 0; }) && plan2.size() < plan.size())
                    plan.swap(plan2);
            }
        }
        winReg = w == 4;
        for (const Ph &p : phs) {
            const uint32_t lo = p.ua & (~p.ua + 1);
            winReg = winReg && p.ua == p.ub && p.ua == (lo | (lo << 1));
        }
        winTu = winReg;  // narrowed per phase below
        if (winW < 0 && !winReg) return true;  // automatic: register tiles only
        // tables: one blob of 32-bit words per pass and the class member
        // lists (rows)
        const uint32_t full = nq >= 32 ? 0xffffffffu : ((1u << nq) - 1);
        // x / d = (x * ceil(2^24 / d)) >> 24 exactly for x < 2^24 / d
        auto magic = [](int d) { return d > 0 ? (uint32_t)((((uint64_t)1 << 24) + d - 1) / d) : 0u; };
        std::vector<uint32_t> blob, rows;
        std::vector<int> kpOp;
        for (const auto &pl : plan) {
            WinPass wp;
            wp.w0 = pl.first;
            const uint32_t W = ((1u << w) - 1) << wp.w0;
            std::vector<std::vector<uint32_t>> byCount(w + 1);
            for (uint32_t x = 0; x < (1u << w); x++) byCount[__builtin_popcount(x)].push_back(x);
            for (int c = 0; c <= w; c++) wp.maxSide = std::max(wp.maxSide, (int)byCount[c].size());
            std::vector<uint32_t> pat, cls, cstart;
            std::vector<int> coff;
            for (int c = 0; c <= w; c++) {
                coff.push_back((int)pat.size());
                for (uint32_t x : byCount[c]) pat.push_back(x << wp.w0);
            }
            coff.push_back(1 << w);
            std::map<uint32_t, int> mA, mB;
            for (uint32_t st : unA) mA[st & ~W] = __builtin_popcount(st & W);
            for (uint32_t st : unB) mB[st & ~W] = __builtin_popcount(st & W);
            wp.nAcls = (int)mA.size();
            wp.nBcls = (int)mB.size();
            wp.rowsOff = (int)rows.size();
            std::vector<int> &cntA = wp.cntA, &cntB = wp.cntB;
            for (int spin = 0; spin < 2; spin++) {
                const int base = (int)rows.size();
                for (const auto &kv : spin ? mB : mA) {
                    cls.push_back(kv.first | ((uint32_t)kv.second << 24));
                    cstart.push_back((uint32_t)((int)rows.size() - base));
                    (spin ? wp.startB : wp.startA).push_back((int)rows.size() - base);
                    (spin ? wp.incB : wp.incA).push_back(kv.second);
                    (spin ? wp.outB : wp.outA).push_back(kv.first);
                    (spin ? cntB : cntA).push_back((int)byCount[kv.second].size());
                    for (uint32_t x : byCount[kv.second]) {
                        const uint32_t st = kv.first | (x << wp.w0);
                        rows.push_back(spin ? (uint32_t)rkB[st] : (uint32_t)rkA[st] * (uint32_t)nb);
                    }
                }
                if (spin == 0) wp.nrowsA = (int)rows.size() - base;
            }
            wp.tileSize.resize((size_t)wp.nAcls * wp.nBcls);
            for (int ia = 0; ia < wp.nAcls; ia++)
                for (int ib = 0; ib < wp.nBcls; ib++) {
                    wp.tileSize[(size_t)ia * wp.nBcls + ib] = cntA[ia] * cntB[ib];
                    wp.maxMembers = std::max(wp.maxMembers, cntA[ia] * cntB[ib]);
                }
            wp.allTiles.resize(wp.tileSize.size());
            std::iota(wp.allTiles.begin(), wp.allTiles.end(), 0);
            // largest first, and one inner-count class after another within a
            // size, so the waves resident at a time run few of the register
            // kernels' per-class code paths
            auto icls = [&](int t) {
                return wp.incA[t / wp.nBcls] * (kWinMaxW + 1) + wp.incB[t % wp.nBcls];
            };
            std::stable_sort(wp.allTiles.begin(), wp.allTiles.end(), [&](int x, int y) {
                return wp.tileSize[x] != wp.tileSize[y] ? wp.tileSize[x] > wp.tileSize[y] : icls(x) < icls(y);
            });
            // phases, sub-class lists, programs, masks
            std::vector<WinPhaseDev> ph;
            std::vector<int4> subIdx;
            std::vector<uint32_t> sub, mask;
            std::vector<Seg> psegs;
            wp.kp0 = (int)kpOp.size();
            for (int i : pl.second) {
                const Ph &p = phs[i];
                WinPhaseDev d;
                d.k0 = p.k0;
                d.nk = p.k1 - p.k0;
                d.progOff = 0;
                d.kp = (int)kpOp.size() - wp.kp0;
                d.ua = p.ua;
                d.ub = p.ub;
                d.spect = full & ~((p.ua << n) | p.ub);
                d.pad = 0;
                for (int k = p.k0; k < p.k1; k++) {
                    kpOp.push_back(k);
                    mask.push_back(hops[k].mask);
                }
                for (int spin = 0; spin < 2; spin++) {
                    const uint32_t u = (spin ? p.ub : p.ua) >> wp.w0;
                    for (int c = 0; c <= w; c++) {
                        const std::vector<uint32_t> &xs = byCount[c];
                        std::map<uint32_t, std::vector<int>> grp;
                        for (int i2 = 0; i2 < (int)xs.size(); i2++) grp[xs[i2] & ~u].push_back(i2);
                        const int off = (int)sub.size();
                        int n2 = 0;
                        for (int pass = 0; pass < 2; pass++)
                            for (const auto &g : grp) {
                                const bool two = g.second.size() == 2;
                                if (two != (pass == 0)) continue;
                                if (two) {
                                    sub.push_back((uint32_t)g.second[0] | ((uint32_t)g.second[1] << 8));
                                    n2++;
                                } else {
                                    const uint32_t sh = (xs[g.second[0]] & u) ? 2u : 1u;
                                    sub.push_back((uint32_t)g.second[0] | (0xffu << 8) | (sh << 16));
                                }
                            }
                        const int ng = (int)grp.size();
                        subIdx.push_back(make_int4(off, ng | (n2 << 16), (int)magic(ng), (int)magic(n2)));
                    }
                }
                ph.push_back(d);
                Seg sg{p.k0, p.k1, -1};
                sg.orbA = 0;
                sg.ua = p.ua;
                sg.ub = p.ub;
                psegs.push_back(sg);
            }
            wp.nph = (int)pl.second.size();
            wp.nkp = (int)kpOp.size() - wp.kp0;
            // orbit programs of the phases (9 types x kOrbMaxOps words each)
            std::vector<uint32_t> prog;
            for (Seg &sg : psegs) {
                sg.progOff = (int)prog.size();
                orbit_program_words(sg, prog);
            }
            for (size_t i = 0; i < ph.size(); i++) ph[i].progOff = psegs[i].progOff;
            // TUPS units: beta single, alpha single, paired double, beta
            // single, alpha single on one pair, each rotating exactly the
            // member pairs of its kind in every orbit type, with an empty
            // sign mask (no spin-orbital between its indices) and so one
            // phase for every pair it rotates
            auto op_phase = [&](const WinPhaseDev &d, int kk) {
                int sg = -1;  // -1 no pair yet, 0 or 1, 2 mixed
                for (int ty = 0; ty < 9; ty++) {
                    const uint32_t w = prog[d.progOff + ty * kOrbMaxOps + kk];
                    for (int r = 0; r < 6; r++)
                        if ((w >> r) & 1u) {
                            const int b = (int)((w >> (8 + r)) & 1u);
                            sg = sg < 0 ? b : (sg == b ? sg : 2);
                        }
                }
                return sg;
            };
            for (size_t i = 0; winTu && i < ph.size(); i++) {
                static const int kind[5] = {0, 1, 2, 0, 1};
                winTu = ph[i].nk == 5;
                for (int kk = 0; winTu && kk < 5; kk++) {
                    const int k = ph[i].k0 + kk;
                    const uint32_t a2 = bitsA(k), b2 = bitsB(k);
                    winTu = (kind[kk] == 0 ? (a2 == 0 && b2 == ph[i].ub)
                             : kind[kk] == 1 ? (a2 == ph[i].ua && b2 == 0) : (a2 == ph[i].ua && b2 == ph[i].ub));
                    winTu = winTu && hops[k].mask == 0 && (op_phase(ph[i], kk) == 0 || op_phase(ph[i], kk) == 1);
                    for (int ta = 0; winTu && ta < 3; ta++)
                        for (int tb = 0; winTu && tb < 3; tb++) {
                            const unsigned want = kind[kk] == 0 ? (tb == 0 ? (1u | (ta == 0 ? 32u : 0u)) : 0u)
                                                : kind[kk] == 1 ? (ta == 0 ? (2u | (tb == 0 ? 16u : 0u)) : 0u)
                                                                : (ta == 0 && tb == 0 ? 4u : 0u);
                            winTu = (prog[ph[i].progOff + (ta * 3 + tb) * kOrbMaxOps + kk] & 0x3fu) == want;
                        }
                }
            }
            // the blob
            std::vector<uint32_t> bl;
            auto put = [&](const void *src, size_t bytes) {
                const size_t o = bl.size();
                bl.resize(o + (bytes + 3) / 4);
                memcpy(bl.data() + o, src, bytes);
                return (int)o;
            };
            put(ph.data(), ph.size() * sizeof(WinPhaseDev));
            while (bl.size() % 4) bl.push_back(0);  // int4 alignment
            wp.oSubIdx = put(subIdx.data(), subIdx.size() * sizeof(int4));
            wp.oSub = put(sub.data(), sub.size() * 4);
            wp.oProg = put(prog.data(), prog.size() * 4);
            wp.oMask = put(mask.data(), mask.size() * 4);
            wp.oPat = put(pat.data(), pat.size() * 4);
            wp.oCOff = put(coff.data(), coff.size() * 4);
            wp.oCls = put(cls.data(), cls.size() * 4);
            wp.oCStart = put(cstart.data(), cstart.size() * 4);
            // the programs by (phase, operator): the 9 type words of one
            // operator together (16 words, one wide scalar load); word 9: the
            // operator's phase as a sign bit (0x80000000) when all its pairs
            // share one (the straight-line TUPS form reads it)
            while (bl.size() % 16) bl.push_back(0);
            wp.oProgT = (int)bl.size();
            for (size_t q = 0; q < ph.size(); q++)
                for (int kk = 0; kk < kOrbMaxOps; kk++)
                    for (int ty = 0; ty < 16; ty++)
                        bl.push_back(ty < 9 ? prog[ph[q].progOff + ty * kOrbMaxOps + kk]
                                            : ty == 9 && kk < ph[q].nk && op_phase(ph[q], kk) == 1 ? 0x80000000u : 0u);
            while (bl.size() % 4) bl.push_back(0);
            wp.blobOff = (int)blob.size();
            wp.nblob = (int)bl.size();
            blob.insert(blob.end(), bl.begin(), bl.end());
            wpass.push_back(std::move(wp));
        }
        if (!upload_vec(d_wblob, blob, err) || !upload_vec(d_wrows, rows, err) || !upload_vec(d_wkp, kpOp, err))
            return false;
        hrows = rows;
        // every tile listed until the reachable lists are computed
        std::vector<int> tl;
        for (WinPass &wp : wpass) {
            wp.tileOff = (int)tl.size();
            wp.ntiles = (int)wp.allTiles.size();
            tl.insert(tl.end(), wp.allTiles.begin(), wp.allTiles.end());
            wp.allOff = wp.tileOff;
        }
        if (!upload_vec(d_wtiles, tl, err) || !upload_vec(d_wall, tl, err)) return false;
        // vector-major tasks (register tiles): the tiles by class of inner
        // counts (5 ca + cb), up to 64 of one class per task.  A task's record
        // block holds, per lane (tile) and word q, q < NA its alpha rows, then
        // NB beta rows, then its alpha and beta outer bits: (NA + NB + 2) x 64
        // words, so a lane finds its members in one coalesced load per word
        if (winReg) {
            std::vector<uint32_t> rec;
            std::vector<int4> tasks;
            for (WinPass &wp : wpass) {
                wp.vRecOff = (int)rec.size();
                wp.vTaskOff = (int)tasks.size();
                std::vector<std::vector<int>> byCls(25);
                for (int t : wp.allTiles) byCls[wp.incA[t / wp.nBcls] * 5 + wp.incB[t % wp.nBcls]].push_back(t);
                const uint32_t *rA = hrows.data() + wp.rowsOff, *rB = rA + wp.nrowsA;
                for (int cl = 0; cl < 25; cl++)
                    for (size_t i = 0; i < byCls[cl].size(); i += 64) {
                        const int cnt = (int)std::min<size_t>(64, byCls[cl].size() - i);
                        const int t0 = byCls[cl][i];
                        const int NA = wp.cntA[t0 / wp.nBcls], NB = wp.cntB[t0 % wp.nBcls];
                        const size_t off = rec.size();
                        tasks.push_back(make_int4(cl, (int)(off - wp.vRecOff), cnt, 0));
                        rec.resize(off + (size_t)(NA + NB + 2) * kLanes, 0u);
                        for (int ln = 0; ln < cnt; ln++) {
                            const int t = byCls[cl][i + ln], ia = t / wp.nBcls, ib = t - ia * wp.nBcls;
                            for (int q = 0; q < NA; q++) rec[off + (size_t)q * kLanes + ln] = rA[wp.startA[ia] + q];
                            for (int q = 0; q < NB; q++)
                                rec[off + (size_t)(NA + q) * kLanes + ln] = rB[wp.startB[ib] + q];
                            rec[off + (size_t)(NA + NB) * kLanes + ln] = wp.outA[ia];
                            rec[off + (size_t)(NA + NB + 1) * kLanes + ln] = wp.outB[ib];
                        }
                    }
                wp.vTasks = (int)tasks.size() - wp.vTaskOff;
            }
            if (!upload_vec(d_wvrec, rec, err) || !upload_vec(d_wvtasks, tasks, err)) return false;
            // staged alpha rows: the beta outer classes by inner count, up to
            // 64 per task; record block per task: NB beta rows, then the beta
            // outer bits, (NB + 1) x 64 words
            std::vector<uint32_t> srec;
            std::vector<int4> stasks;
            winVsOK = true;
            for (WinPass &wp : wpass) {
                wp.vsRecOff = (int)srec.size();
                wp.vsTaskOff = (int)stasks.size();
                const uint32_t *rB = hrows.data() + wp.rowsOff + wp.nrowsA;
                for (int cb = 0; cb <= 4; cb++) {
                    std::vector<int> jbs;
                    for (int jb = 0; jb < wp.nBcls; jb++)
                        if (wp.incB[jb] == cb) jbs.push_back(jb);
                    for (size_t i = 0; i < jbs.size(); i += 64) {
                        const int cnt = (int)std::min<size_t>(64, jbs.size() - i);
                        const int NB = wp.cntB[jbs[i]];
                        const size_t off = srec.size();
                        stasks.push_back(make_int4(cb, (int)(off - wp.vsRecOff), cnt, 0));
                        srec.resize(off + (size_t)(NB + 1) * kLanes, 0u);
                        for (int ln = 0; ln < cnt; ln++) {
                            const int jb = jbs[i + ln];
                            for (int q = 0; q < NB; q++) srec[off + (size_t)q * kLanes + ln] = rB[wp.startB[jb] + q];
                            srec[off + (size_t)NB * kLanes + ln] = wp.outB[jb];
                        }
                    }
                }
                wp.vsTasks = (int)stasks.size() - wp.vsTaskOff;
                wp.vsWavesF = std::max(1, std::min(16, wp.vsTasks));
                wp.vsLdsF = (size_t)kWinVsRows * nb * sizeof(double);
                // the adjoint: as many waves (at most 8) as leave room for their bracket rows
                const size_t perWave = (size_t)wp.nkp * kLanes * sizeof(double);
                const long long room = (long long)ldsDevice - 2 * (long long)wp.vsLdsF;
                wp.vsWavesA = (int)std::min<long long>(std::min(8, wp.vsTasks), room > 0 ? room / (long long)perWave : 0);
                wp.vsLdsA = 2 * wp.vsLdsF + (size_t)std::max(1, wp.vsWavesA) * perWave;
                winVsOK = winVsOK && wp.vsWavesA >= 1 && wp.vsLdsA <= ldsDevice;
            }
            if (!upload_vec(d_wvsrec, srec, err) || !upload_vec(d_wvstasks, stasks, err)) return false;
        }
        // grids: as many blocks as the chip holds at once
        std::vector<int> lrows(M, 0);
        wslots = 1;
        for (WinPass &wp : wpass) {
            for (int dir = 0; dir < 2; dir++) {
                const bool br = dir == 1;
                const size_t lds = win_lds(wp, br);
                if (lds > ldsDevice) {
                    wpass.clear();
                    winReg = false;
                    return true;  // a tile does not fit: no window plan
                }
                const int nw = win_waves(br);
                int nbk = 0;
                for (int bc = 0; bc < (br ? 1 : 2); bc++) {
                    const void *fn = winReg ? winr_fn(br, bc == 1) : win_fn(br, nw, bc == 1);
                    HCHK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)std::max<size_t>(lds, 1)));
                }
                if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                        &nbk, winReg ? winr_fn(br, false) : win_fn(br, nw, false), nw * kLanes, lds) != hipSuccess ||
                    nbk < 1)
                    nbk = 1;
                // register tiles: one wave per tile
                const long long work = winReg ? ((long long)wp.allTiles.size() + nw - 1) / nw : (long long)wp.allTiles.size();
                const long long g = std::min<long long>((long long)nbk * numCU, work);
                (br ? wp.gridA : wp.gridF) = (int)std::max<long long>(1, g);
            }
            for (int q = 0; q < wp.nkp; q++) lrows[kpOp[wp.kp0 + q]] = wp.gridA;
            wslots = std::max(wslots, wp.gridA);
        }
        if (winReg && winTu)
            for (WinPass &wp : wpass) {
                // Hessian tangent passes: 2 waves per block, at most 4 x CUs blocks (its partial rows)
                const size_t lds = (size_t)2 * wp.nkp * (kLanes + 1) * sizeof(double);
                HCHK(hipFuncSetAttribute((const void *)k_wint, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
                int nbk = 0;
                if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nbk, (const void *)k_wint, 2 * kLanes, lds) !=
                        hipSuccess ||
                    nbk < 1)
                    nbk = 1;
                const long long g = std::min<long long>(std::min<long long>((long long)nbk * numCU, 4LL * numCU),
                                                        ((long long)wp.allTiles.size() + 1) / 2);
                wp.gridT = (int)std::max<long long>(1, g);
            }
        if (winReg)
            for (WinPass &wp : wpass)
                for (int br = 0; br < 2; br++) {
                    const void *fn = winv_fn(br == 1);
                    const size_t lds = br ? (size_t)4 * wp.nkp * kLanes * sizeof(double) : 0;
                    HCHK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)std::max<size_t>(lds, 1)));
                    int nbk = 0;
                    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nbk, fn, 4 * kLanes, lds) != hipSuccess || nbk < 1)
                        nbk = 1;
                    long long g = std::min<long long>((long long)nbk * numCU, (wp.vTasks + 3) / 4);
                    if (br) g = std::min<long long>(g, slots);  // one partial row per block
                    (br ? wp.gridVA : wp.gridVF) = (int)std::max<long long>(1, g);
                }
        if (winReg && winVsOK)
            for (WinPass &wp : wpass)
                for (int br = 0; br < 2; br++) {
                    const void *fn = winvs_fn(br == 1);
                    const size_t lds = br ? wp.vsLdsA : wp.vsLdsF;
                    const int nw = br ? wp.vsWavesA : wp.vsWavesF;
                    HCHK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
                    int nbk = 0;
                    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nbk, fn, nw * kLanes, lds) != hipSuccess || nbk < 1)
                        nbk = 1;
                    long long g = std::min<long long>((long long)nbk * numCU, wp.nAcls);
                    if (br) g = std::min<long long>(g, slots);
                    (br ? wp.gridSA : wp.gridSF) = (int)std::max<long long>(1, g);
                }
        for (int k = 0; k < M; k++)
            if (lrows[k] == 0) {
                err = "window plan misses an operator";
                return false;
            }
        if (!d_lrows.ensure(std::max(M, 1), err)) return false;
        HCHK(hipMemcpy(d_lrows.p, lrows.data(), M * sizeof(int), hipMemcpyHostToDevice));
        lslots = std::max(lslots, wslots);
        winOn = true;
        supDirty = hessDirty = true;
        return true;
    }

    // The Hessian's tangent passes (k_wint): per pass every tile, the ones
    // the start state and the passes before can reach first (wp.nReach of
    // them), as compute_support_win finds them; option 18 = 0 counts every
    // tile as reached
    bool hessian_tiles(std::string &err)
    {
        if (!hessDirty) return true;
        const bool lists_on = reachOpt && (long long)hstart.size() == N;
        std::vector<uint8_t> S(N, (uint8_t)1);
        if (lists_on)
            for (long long i = 0; i < N; i++) S[i] = hstart[i] != 0.0;
        std::vector<int> tl, rest;
        std::vector<long long> mem;
        for (WinPass &wp : wpass) {
            wp.allOff = (int)tl.size();
            rest.clear();
            const uint32_t *rA = hrows.data() + wp.rowsOff, *rB = rA + wp.nrowsA;
            for (int t : wp.allTiles) {
                const int ia = t / wp.nBcls, ib = t - ia * wp.nBcls;
                mem.clear();
                bool any = false;
                for (int i = 0; i < wp.cntA[ia]; i++)
                    for (int j = 0; j < wp.cntB[ib]; j++) {
                        const long long d = (long long)rA[wp.startA[ia] + i] + rB[wp.startB[ib] + j];
                        mem.push_back(d);
                        any = any || S[d];
                    }
                if (!any) {
                    rest.push_back(t);
                    continue;
                }
                tl.push_back(t);
                for (long long d : mem) S[d] = 1;
            }
            wp.nReach = (int)tl.size() - wp.allOff;
            tl.insert(tl.end(), rest.begin(), rest.end());
        }
        if (!upload_vec(d_wall, tl, err)) return false;
        hessDirty = false;
        return true;
    }

    WinArgs winargs(const WinPass &wp) const
    {
        WinArgs a;
        a.n = n;
        a.w = std::min(std::min(winW < 0 ? 4 : winW, kWinMaxW), n);
        a.blob = d_wblob.p + wp.blobOff;
        a.w0 = wp.w0;
        a.oSubIdx = wp.oSubIdx;
        a.oSub = wp.oSub;
        a.oProg = wp.oProg;
        a.oMask = wp.oMask;
        a.oPat = wp.oPat;
        a.oCOff = wp.oCOff;
        a.oCls = wp.oCls;
        a.oCStart = wp.oCStart;
        a.oProgT = wp.oProgT;
        a.rows = d_wrows.p + wp.rowsOff;
        a.nrowsA = wp.nrowsA;
        a.tiles = d_wtiles.p + wp.tileOff;
        a.ntiles = wp.ntiles;
        a.nAcls = wp.nAcls;
        a.nBcls = wp.nBcls;
        a.nph = wp.nph;
        a.nkp = wp.nkp;
        a.kpOp = d_wkp.p + wp.kp0;
        a.maxMembers = wp.maxMembers;
        return a;
    }

    size_t win_lds(const WinPass &wp, bool br) const
    {
        if (winReg) return br ? (size_t)2 * wp.nkp * kLanes * sizeof(double) : (size_t)wp.nkp * kLanes * 16;
        return win_lds_bytes(br, winWaves, wp.maxMembers, wp.nkp);
    }

    // Reachable tiles of the window passes: S_0 = the start state's nonzero
    // determinants; pass p processes the tiles meeting S_p and S_{p+1} = every
    // member of those tiles.  A tile outside holds exact zeros before and
    // after the pass.  The adjoint pass p takes the same tiles: S_{p+1} is the
    // union of them, and the lambda of a determinant outside S_{p+1} never
    // meets a nonzero state in the passes before it.
    bool compute_support_win(bool ck, std::string &err)
    {
        supDirty = false;
        supCk = ck;
        supGen++;
        const bool lists_on = reachOpt && (long long)hstart.size() == N;
        std::vector<uint8_t> S(N, (uint8_t)1);
        if (lists_on)
            for (long long i = 0; i < N; i++) S[i] = hstart[i] != 0.0;
        ampF = ampAL = ampAP = 0;
        std::vector<int> tl;
        std::vector<long long> mem;
        for (WinPass &wp : wpass) {
            wp.tileOff = (int)tl.size();
            const uint32_t *rA = hrows.data() + wp.rowsOff, *rB = rA + wp.nrowsA;
            for (int t : wp.allTiles) {
                const int ia = t / wp.nBcls, ib = t - ia * wp.nBcls;
                mem.clear();
                bool any = false;
                for (int i = 0; i < wp.cntA[ia]; i++)
                    for (int j = 0; j < wp.cntB[ib]; j++) {
                        const long long d = (long long)rA[wp.startA[ia] + i] + rB[wp.startB[ib] + j];
                        mem.push_back(d);
                        any = any || S[d];
                    }
                if (!any) continue;
                tl.push_back(t);
                for (long long d : mem) S[d] = 1;
                ampF += (long long)mem.size();
            }
            wp.ntiles = (int)tl.size() - wp.tileOff;
        }
        ampAL = ampAP = ampF;
        if (!upload_vec(d_wtiles, tl, err)) return false;
        return true;
    }

    bool sweep_forward_win(double *v, int nch, std::string &err)
    {
        const long long cstride = N * kLanes;
        const dim3 blk(winWaves * kLanes);
        for (size_t p = 0; p < wpass.size(); p++) {
            const WinPass &wp = wpass[p];
            const WinArgs a = winargs(wp);
            const size_t lds = win_lds(wp, false);
            const double *in = ckOn ? ck_state((int)p, nch) : v;
            double *out = ckOn ? ck_state((int)p + 1, nch) : v;
            const bool bc = p == 0 && ckOn;  // pass 0 reads the start vector
            if (bc) in = d_start.p;
            const int mk = mark_begin(1);
            if (winReg) {
                const dim3 grid((unsigned)std::max(1, std::min(wp.gridF, (wp.ntiles + 3) / 4)), nch);
                if (winTu) {
                    if (bc) hipLaunchKernelGGL((k_winr<false, true, 4, true>), grid, dim3(4 * kLanes), lds, stream, a,
                                               in, out, nullptr, cstride, bl_cs.p, M, nullptr, 0, nullptr);
                    else hipLaunchKernelGGL((k_winr<false, false, 4, true>), grid, dim3(4 * kLanes), lds, stream, a,
                                            in, out, nullptr, cstride, bl_cs.p, M, nullptr, 0, nullptr);
                } else {
                    if (bc) hipLaunchKernelGGL((k_winr<false, true, 4, false>), grid, dim3(4 * kLanes), lds, stream, a,
                                               in, out, nullptr, cstride, bl_cs.p, M, nullptr, 0, nullptr);
                    else hipLaunchKernelGGL((k_winr<false, false, 4, false>), grid, dim3(4 * kLanes), lds, stream, a,
                                            in, out, nullptr, cstride, bl_cs.p, M, nullptr, 0, nullptr);
                }
                mark_end(mk);
                LCHK();
                continue;
            }
            const dim3 grid((unsigned)std::max(1, std::min(wp.gridF, wp.ntiles)), nch);
            if (winWaves == 8) {
                if (bc) hipLaunchKernelGGL((k_win<false, true, 8>), grid, blk, lds, stream, a, in, out, nullptr,
                                           cstride, bl_cs.p, M, nullptr, 0);
                else hipLaunchKernelGGL((k_win<false, false, 8>), grid, blk, lds, stream, a, in, out, nullptr,
                                        cstride, bl_cs.p, M, nullptr, 0);
            } else {
                if (bc) hipLaunchKernelGGL((k_win<false, true, 4>), grid, blk, lds, stream, a, in, out, nullptr,
                                           cstride, bl_cs.p, M, nullptr, 0);
                else hipLaunchKernelGGL((k_win<false, false, 4>), grid, blk, lds, stream, a, in, out, nullptr,
                                        cstride, bl_cs.p, M, nullptr, 0);
            }
            mark_end(mk);
            LCHK();
        }
        return true;
    }

    // every block writes its bracket rows, tiles or not (wp.gridA rows per
    // operator).  epart (register tiles): the energies' partial rows from the
    // first adjoint pass (win_erows() per chunk)
    bool sweep_reverse_win(double *phi, double *lam, int nch, std::string &err, double *epart = nullptr)
    {
        const long long cstride = N * kLanes;
        const dim3 blk(winWaves * kLanes);
        for (int p = (int)wpass.size() - 1; p >= 0; p--) {
            const WinPass &wp = wpass[p];
            const WinArgs a = winargs(wp);
            const size_t lds = win_lds(wp, true);
            const double *in = ckOn ? ck_state(p + 1, nch) : phi;
            double *out = ckOn ? nullptr : phi;
            const dim3 grid((unsigned)wp.gridA, nch);
            const int mk = mark_begin(2);
            double *ep = p == (int)wpass.size() - 1 ? epart : nullptr;
            if (winReg && winTu)
                hipLaunchKernelGGL((k_winr<true, false, 2, true>), grid, dim3(2 * kLanes), lds, stream, a, in, out,
                                   lam, cstride, bl_csneg.p, M, bl_partial.p, lslots, ep);
            else if (winReg)
                hipLaunchKernelGGL((k_winr<true, false, 2, false>), grid, dim3(2 * kLanes), lds, stream, a, in, out,
                                   lam, cstride, bl_csneg.p, M, bl_partial.p, lslots, ep);
            else if (winWaves == 8)
                hipLaunchKernelGGL((k_win<true, false, 8>), grid, blk, lds, stream, a, in, out, lam, cstride,
                                   bl_csneg.p, M, bl_partial.p, lslots);
            else
                hipLaunchKernelGGL((k_win<true, false, 4>), grid, blk, lds, stream, a, in, out, lam, cstride,
                                   bl_csneg.p, M, bl_partial.p, lslots);
            mark_end(mk);
            LCHK();
        }
        return true;
    }

    bool build_lane_list(LaneList &L, int maxG, std::string &err)
    {
        std::vector<Seg> &lsegs = L.segs;
        lsegs.clear();
        if (mode != kSz || sweepKind == kSweepTile) {
            lsegs = segs;
            for (Seg &sg : lsegs) sg.two = sg.three = 0;
            return build_orbit_programs(lsegs, L.lprog, err);
        }
        const uint32_t amask = (1u << n) - 1;
        auto bitsA = [&](int kk) { return ((hops[kk].cre | hops[kk].ann) >> n) & amask; };
        auto bitsB = [&](int kk) { return (hops[kk].cre | hops[kk].ann) & amask; };
        const int maxOps = std::min(kOrb2MaxOps, fuseDepth > 0 ? fuseDepth : kOrb2MaxOps);
        const int maxOps3 = std::min(kOrb3MaxOps, fuseDepth > 0 ? fuseDepth : kOrb3MaxOps);
        std::vector<unsigned long long> prog2, prog3;
        // a one-group run from k0: operators whose alpha bits are empty or one
        // fixed 2-bit group, and likewise beta (a TUPS unit)
        auto run = [&](int k0, uint32_t &ua, uint32_t &ub) {
            ua = ub = 0;
            int k1 = k0;
            while (k1 < M && k1 - k0 < std::min(kOrbMaxOps, maxOps)) {
                const uint32_t a = bitsA(k1), b = bitsB(k1);
                if ((a | b) == 0) break;
                if (a && (__builtin_popcount(a) != 2 || (ua && ua != a))) break;
                if (b && (__builtin_popcount(b) != 2 || (ub && ub != b))) break;
                if (a) ua = a;
                if (b) ub = b;
                k1++;
            }
            return k1;
        };
        auto pat = [](uint32_t grp, int shp, int m) -> uint32_t {
            if (shp == 1 || grp == 0) return 0u;
            if (shp == 2) return grp;
            const uint32_t lo = grp & (~grp + 1);
            return m ? (grp ^ lo) : lo;
        };
        static const int PP[6] = {0, 0, 0, 1, 1, 2}, QQ[6] = {1, 2, 3, 2, 3, 3};
        // programs of a two-group segment [k, k2) split at k1: false if some
        // operator's pair pattern differs between its sub-orbits
        auto two_group_prog = [&](const Seg &sg, int k, int k1, int k2, std::vector<unsigned long long> &out,
                                  int K) {
            out.clear();
            for (int type = 0; type < 81; type++) {
                const int sa1 = type % 3, sa2 = (type / 3) % 3, sb1 = (type / 9) % 3, sb2 = (type / 27) % 3;
                auto valid = [&](int slot) {
                    const int ma1 = slot & 1, ma2 = (slot >> 1) & 1, mb1 = (slot >> 2) & 1, mb2 = (slot >> 3) & 1;
                    const bool twoA1 = sa1 == 0 && sg.grA[0], twoA2 = sa2 == 0 && sg.grA[1];
                    const bool twoB1 = sb1 == 0 && sg.grB[0], twoB2 = sb2 == 0 && sg.grB[1];
                    return (ma1 == 0 || twoA1) && (ma2 == 0 || twoA2) && (mb1 == 0 || twoB1) && (mb2 == 0 || twoB2);
                };
                auto str = [&](int slot) {
                    const int ma1 = slot & 1, ma2 = (slot >> 1) & 1, mb1 = (slot >> 2) & 1, mb2 = (slot >> 3) & 1;
                    const uint32_t a = pat(sg.grA[0], sa1, ma1) | pat(sg.grA[1], sa2, ma2);
                    const uint32_t b = pat(sg.grB[0], sb1, mb1) | pat(sg.grB[1], sb2, mb2);
                    return (a << n) | b;
                };
                for (int kk = 0; kk < K; kk++) {
                    unsigned long long w = 0;
                    if (k + kk < k2) {
                        const DevOp &o = hops[k + kk];
                        const int g = (k + kk < k1) ? 0 : 1, h = g;
                        auto slotc = [&](int c, int i, int j) {
                            const int oa = c & 1, ob = c >> 1;
                            return ((i << g) | (oa << (1 - g))) + 4 * ((j << h) | (ob << (1 - h)));
                        };
                        unsigned long long words[4] = {0, 0, 0, 0};
                        for (int c = 0; c < 4; c++)
                            for (int r = 0; r < 6; r++) {
                                const int sp = slotc(c, PP[r] >> 1, PP[r] & 1), sq = slotc(c, QQ[r] >> 1, QQ[r] & 1);
                                if (!valid(sp) || !valid(sq)) continue;
                                uint32_t b2 = 0;
                                const int sgn = host_pair_sign(o, str(sp), b2);
                                if (sgn == 0 || b2 != str(sq)) continue;
                                words[c] |= 1ull << r;
                                if (sgn < 0) words[c] |= 1ull << (6 + r);
                            }
                        // one word for all four sub-orbits: every sub-orbit has it, or
                        // has both members of each of its pairs absent (absent amplitudes
                        // are zero in registers and never stored)
                        unsigned long long m = 0;
                        for (int c = 0; c < 4; c++) m |= words[c];
                        for (int c = 0; c < 4; c++) {
                            if (words[c] == m) continue;
                            for (int r = 0; r < 6; r++) {
                                if (!((m >> r) & 1)) continue;
                                const int sp = slotc(c, PP[r] >> 1, PP[r] & 1), sq = slotc(c, QQ[r] >> 1, QQ[r] & 1);
                                if (valid(sp) || valid(sq)) return false;
                            }
                        }
                        w = m;
                    }
                    out.push_back(w);
                }
            }
            return true;
        };
        // run-2 words of a three-group segment per group-3 type (shapeA3 + 3
        // shapeB3) over the four group-3 slots (a3 + 2 b3): false if a slot is
        // in two pairs, the pairs differ in their slot xor, or something acts
        // where both group-3 classes have one member (k_orb3v skips run 2 there)
        auto three_group_prog3 = [&](const Seg &sg, int k, int k2, int k3, std::vector<unsigned long long> &out) {
            out.clear();
            const uint32_t g3a = sg.grA[2], g3b = sg.grB[2];
            for (int type3 = 0; type3 < 9; type3++) {
                const int sa = type3 % 3, sb = type3 / 3;
                auto valid = [&](int slot) {
                    return ((slot & 1) == 0 || (sa == 0 && g3a)) && ((slot >> 1) == 0 || (sb == 0 && g3b));
                };
                auto str = [&](int slot) { return (pat(g3a, sa, slot & 1) << n) | pat(g3b, sb, slot >> 1); };
                for (int kk = 0; kk < kOrb3MaxOps; kk++) {
                    unsigned long long w = 0;
                    if (k + kk >= k2 && k + kk < k3) {
                        const DevOp &o = hops[k + kk];
                        int used = 0, dx = -1;
                        for (int r = 0; r < 6; r++) {
                            const int p = PP[r], q = QQ[r];
                            if (!valid(p) || !valid(q)) continue;
                            uint32_t b2 = 0;
                            const int sgn = host_pair_sign(o, str(p), b2);
                            if (sgn == 0 || b2 != str(q)) continue;
                            if (((used >> p) & 1) || ((used >> q) & 1)) return false;
                            if (dx >= 0 && dx != (p ^ q)) return false;
                            used |= (1 << p) | (1 << q);
                            dx = p ^ q;
                            w |= 1ull << r;
                            if (sgn < 0) w |= 1ull << (6 + r);
                        }
                        if (w && sa != 0 && sb != 0) return false;
                    }
                    out.push_back(w);
                }
            }
            return true;
        };
        std::vector<unsigned long long> words, words3;
        int k = 0;
        while (k < M) {
            uint32_t ua = 0, ub = 0;
            const int k1 = run(k, ua, ub);
            if (k1 == k) {
                // not an orbit operator: one generic pass
                Seg sg{k, k + 1, -1};
                lsegs.push_back(sg);
                k++;
                continue;
            }
            if (maxG >= 3 && k1 < M) {
                uint32_t ua2 = 0, ub2 = 0, ua3 = 0, ub3 = 0;
                const int k2 = run(k1, ua2, ub2);
                const int k3 = (k2 > k1 && k2 < M) ? run(k2, ua3, ub3) : k2;
                const uint32_t A12 = ua | ua2, B12 = ub | ub2;
                if (k2 > k1 && k3 > k2 && k3 - k <= maxOps3 && !(ua & ua2) && !(ub & ub2) && !(A12 & ua3) &&
                    !(B12 & ub3)) {
                    // no operator's sign mask may cover another group's orbitals:
                    // the run-0/1 words hold for every group-3 member and the run-2
                    // words for every group-1/2 member
                    bool masks = true;
                    for (int kk = k; kk < k3; kk++) {
                        const uint32_t other = kk < k2 ? ((ua3 << n) | ub3) : ((A12 << n) | B12);
                        if (hops[kk].mask & other) masks = false;
                    }
                    Seg sg{k, k3, -1};
                    sg.two = 1;
                    sg.three = 1;
                    sg.grA[0] = ua;
                    sg.grA[1] = ua2;
                    sg.grA[2] = ua3;
                    sg.grB[0] = ub;
                    sg.grB[1] = ub2;
                    sg.grB[2] = ub3;
                    if (masks && two_group_prog(sg, k, k1, k2, words, kOrb3MaxOps) &&
                        three_group_prog3(sg, k, k2, k3, words3)) {
                        sg.gA = groupset3_for(ua, ua2, ua3, false, err);
                        sg.gB = groupset3_for(ub, ub2, ub3, true, err);
                        if (sg.gA < 0 || sg.gB < 0) return false;
                        const GroupSet &GA = *gsets[sg.gA], &GB = *gsets[sg.gB];
                        const long long a2 = GA.n2, a1 = GA.ncls - GA.n2, b2 = GB.n2, b1 = GB.ncls - GB.n2;
                        sg.orbits = 4 * a2 * b2 + 2 * a2 * b1 + 2 * a1 * b2 + a1 * b1;  // work items
                        sg.prog2Off = (int)prog2.size();
                        prog2.insert(prog2.end(), words.begin(), words.end());
                        sg.prog3Off = (int)prog3.size();
                        prog3.insert(prog3.end(), words3.begin(), words3.end());
                        sg.progOff = k1 - k;
                        sg.split2 = k2 - k;
                        lsegs.push_back(sg);
                        k = k3;
                        continue;
                    }
                }
            }
            if (maxG >= 2 && k1 < M) {
                uint32_t ua2 = 0, ub2 = 0;
                const int k2 = run(k1, ua2, ub2);
                if (k2 > k1 && k2 - k <= std::min(kOrb2MaxOps, maxOps) && !(ua & ua2) && !(ub & ub2)) {
                    Seg sg{k, k2, -1};
                    sg.two = 1;
                    sg.grA[0] = ua;
                    sg.grA[1] = ua2;
                    sg.grB[0] = ub;
                    sg.grB[1] = ub2;
                    if (two_group_prog(sg, k, k1, k2, words, kOrb2MaxOps)) {
                        sg.gA = groupset_for(ua, ua2, false, err);
                        sg.gB = groupset_for(ub, ub2, true, err);
                        if (sg.gA < 0 || sg.gB < 0) return false;
                        sg.orbits = (long long)gsets[sg.gA]->ncls * gsets[sg.gB]->ncls;
                        sg.prog2Off = (int)prog2.size();
                        prog2.insert(prog2.end(), words.begin(), words.end());
                        sg.progOff = k1 - k;  // the split: operators on the first groups
                        lsegs.push_back(sg);
                        k = k2;
                        continue;
                    }
                }
            }
            Seg sg{k, k1, -1};
            sg.ua = ua;
            sg.ub = ub;
            sg.orbA = tileset_for(ua, err, false);
            sg.orbB = tileset_for(ub, err, true);
            if (sg.orbA < 0 || sg.orbB < 0) return false;
            sg.orbits = (long long)tilesets[sg.orbA]->ntiles * tilesets[sg.orbB]->ntiles;
            lsegs.push_back(sg);
            k = k1;
        }
        if (prog2.empty()) prog2.push_back(0);
        if (prog3.empty()) prog3.push_back(0);
        if (!L.prog2.ensure(prog2.size(), err) || !L.prog3.ensure(prog3.size(), err)) return false;
        HCHK(hipMemcpy(L.prog2.p, prog2.data(), prog2.size() * sizeof(unsigned long long), hipMemcpyHostToDevice));
        HCHK(hipMemcpy(L.prog3.p, prog3.data(), prog3.size() * sizeof(unsigned long long), hipMemcpyHostToDevice));
        // one-group segments of the lane list get their own programs
        std::vector<uint32_t> save(lsegs.size());
        for (size_t i = 0; i < lsegs.size(); i++) save[i] = (uint32_t)lsegs[i].progOff;
        if (!build_orbit_programs(lsegs, L.lprog, err)) return false;
        for (size_t i = 0; i < lsegs.size(); i++)
            if (lsegs[i].two) lsegs[i].progOff = (int)save[i];
        return true;
    }

    // Tables of the two-half lane sigma (k_sigh_sparse): beta strings split
    // by the occupation of orbital kSigHOrb (half A occupied), slot = position
    // within the half; target slots A (padded to 4) then B; per target its
    // A-phase entries then its B-phase entries, then rank, own slot offset,
    // string.  Entry coefficient slots as k_beta_table's (r occupied, q empty
    // in the target string; sign half when an odd number of occupied
    // orbitals lies between them).
    bool build_sigh_tables(std::string &err)
    {
        const int kd = __builtin_popcount(unB[0]);
        const int S = (kd - 1) * (n - kd), X = n - kd;
        const uint32_t pbit = 1u << kSigHOrb, full = (1u << n) - 1;
        std::vector<int> hl(2 * kSigHMax, 0), slotOf(nb);
        int nA = 0, nB = 0;
        for (int b = 0; b < nb; b++) {
            if (unB[b] & pbit) {
                slotOf[b] = nA;
                hl[nA++] = b;
            } else {
                slotOf[b] = nB;
                hl[kSigHMax + nB++] = b;
            }
        }
        if (nA > kSigHMax || nB > kSigHMax) {
            err = "two-half sigma: a half exceeds the LDS buffer";
            return false;
        }
        const int TA = (nA + 3) / 4 * 4, T = TA + (nB + 3) / 4 * 4;
        const uint32_t pad = (uint32_t)(2 * kSigNN * 8) << 17;
        std::vector<uint32_t> tab((size_t)T * kSigWNtp, pad);
        auto parity = [](uint32_t st, int r, int q) {
            const int lo = std::min(r, q), hi = std::max(r, q);
            const uint32_t msk = ((1u << hi) - 1) & ~((1u << (lo + 1)) - 1);
            return __builtin_popcount(st & msk) & 1;
        };
        for (int t = 0; t < T; t++) {
            uint32_t *row = &tab[(size_t)t * kSigWNtp];
            const int b = t < TA ? (t < nA ? hl[t] : -1) : (t - TA < nB ? hl[kSigHMax + t - TA] : -1);
            row[kSigWNtp - 3] = 0xffffffffu;
            row[kSigWNtp - 2] = 0;
            row[kSigWNtp - 1] = 0;
            if (b < 0) continue;
            const uint32_t st = unB[b];
            const bool isA = (st & pbit) != 0;
            std::vector<uint32_t> same, cross;
            for (uint32_t o = st; o; o &= o - 1)
                for (uint32_t wq = ~st & full; wq; wq &= wq - 1) {
                    const int r = __builtin_ctz(o), q = __builtin_ctz(wq);
                    const int J = rkB[st ^ (1u << r) ^ (1u << q)];
                    const int ecf = r * n + q + (parity(st, r, q) ? kSigNN : 0);
                    const uint32_t ent = (uint32_t)(slotOf[J] * kSigWVC * 8) | ((uint32_t)(ecf * 8) << 17);
                    (((unB[J] & pbit) != 0) == isA ? same : cross).push_back(ent);
                }
            if ((int)same.size() != S || (int)cross.size() != X) {
                err = "two-half sigma: unexpected entry counts";
                return false;
            }
            int k = 0;
            for (uint32_t ent : isA ? same : cross) row[k++] = ent;
            for (uint32_t ent : isA ? cross : same) row[k++] = ent;
            row[kSigWNtp - 3] = (uint32_t)b;
            row[kSigWNtp - 2] = (uint32_t)(slotOf[b] * kSigWVC * 8);
            row[kSigWNtp - 1] = st;
        }
        if (!d_tsh.ensure(tab.size(), err) || !d_hl.ensure(hl.size(), err)) return false;
        HCHK(hipMemcpy(d_tsh.p, tab.data(), tab.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
        HCHK(hipMemcpy(d_hl.p, hl.data(), hl.size() * sizeof(int), hipMemcpyHostToDevice));
        HCHK(hipFuncSetAttribute((const void *)sigh_kernel(S, X), hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)kSigHLds));
        sighS = S;
        sighX = X;
        sighNA = nA;
        sighNB = nB;
        sighTA = TA;
        sighT = T;
        return true;
    }

    Orb2Args orb2args(const LaneList &L, const Seg &sg, bool reverse) const
    {
        const GroupSet &A = *gsets[sg.gA], &B = *gsets[sg.gB];
        Orb2Args a;
        a.n = n;
        a.nb = nb;
        a.aRanks = A.ranks.p;
        a.bRanks = B.ranks.p;
        a.aInfo = A.info.p;
        a.bInfo = B.info.p;
        a.nAc = A.ncls;
        a.nBc = B.ncls;
        a.nAm = A.nm;
        a.nBm = B.nm;
        a.ops = d_ops.p;
        a.k0 = sg.k0;
        a.k1 = sg.k1;
        a.reverse = reverse ? 1 : 0;
        a.prog = L.prog2.p + sg.prog2Off;
        a.split = sg.progOff;
        a.list = a.llist = nullptr;
        a.nlist = a.nllist = 0;
        if (!reverse && sg.supOff >= 0) {
            a.list = d_suplist.p + sg.supOff;
            a.llist = d_suplist.p + sg.supLOff;
            a.nlist = sg.supN;
            a.nllist = sg.supLN;
        }
        if (reverse && ckOn && sg.supOff >= 0) {
            a.list = d_suplist.p + sg.supOff;
            a.nlist = sg.supN;
        }
        return a;
    }

    Orb3Args orb3args(const LaneList &L, const Seg &sg) const
    {
        const GroupSet &A = *gsets[sg.gA], &B = *gsets[sg.gB];
        Orb3Args a;
        a.n = n;
        a.nb = nb;
        a.aRanks = A.ranks.p;
        a.bRanks = B.ranks.p;
        a.aInfo = A.info.p;
        a.bInfo = B.info.p;
        a.nA = A.ncls;
        a.nB = B.ncls;
        a.nA2 = A.n2;
        a.nB2 = B.n2;
        const long long a1 = A.ncls - A.n2, b1 = B.ncls - B.n2;
        a.w[0] = 4LL * A.n2 * B.n2;
        a.w[1] = a.w[0] + 2LL * A.n2 * b1;
        a.w[2] = a.w[1] + 2LL * a1 * B.n2;
        a.w[3] = a.w[2] + a1 * b1;
        a.ops = d_ops.p;
        a.k0 = sg.k0;
        a.k1 = sg.k1;
        a.prog = L.prog2.p + sg.prog2Off;
        a.prog3 = L.prog3.p + sg.prog3Off;
        a.s1 = sg.progOff;
        a.s2 = sg.split2;
        return a;
    }

    int generation = 0;  // bumped whenever captured sequences become stale

    void drop_graphs()
    {
        generation++;
        if (egGraph) (void)hipGraphExecDestroy(egGraph);
        if (bGraph) (void)hipGraphExecDestroy(bGraph);
        egGraph = bGraph = nullptr;
    }

    // HIP-event timing of kernel classes (see Engine::set_timing)
    bool timingOn = false;
    struct Mark {
        int cls;
        hipEvent_t a, b;
    };
    std::vector<Mark> marks;
    std::vector<hipEvent_t> evpool;
    static constexpr int kClasses = 4;  // 0 Hamiltonian (sparse part), 1 forward, 2 adjoint, 3 dense same-spin GEMMs
    double tms[kClasses] = {0, 0, 0, 0};
    long long tcnt[kClasses] = {0, 0, 0, 0};

    hipEvent_t take_event()
    {
        if (!evpool.empty()) {
            hipEvent_t e = evpool.back();
            evpool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        (void)hipEventCreate(&e);
        return e;
    }
    int mark_begin(int cls)
    {
        if (!timingOn) return -1;
        Mark m{cls, take_event(), take_event()};
        (void)hipEventRecord(m.a, stream);
        marks.push_back(m);
        return (int)marks.size() - 1;
    }
    void mark_end(int id)
    {
        if (id >= 0) (void)hipEventRecord(marks[id].b, stream);
    }
    // after a stream synchronisation
    void collect_marks()
    {
        for (auto &m : marks) {
            float ms = 0;
            if (hipEventElapsedTime(&ms, m.a, m.b) == hipSuccess) {
                tms[m.cls] += ms;
                tcnt[m.cls] += 1;
            }
            evpool.push_back(m.a);
            evpool.push_back(m.b);
        }
        marks.clear();
    }

    // energy partials per vector: one per row block (rows) or per alpha
    // string (matrix-free); buffers are sized for the larger of the two
    int eslots() const { return hkind == 3 ? na : (int)((N + kBlock - 1) / kBlock); }
    int eslots_alloc() const { return std::max(na, (int)((N + kBlock - 1) / kBlock)); }

    MfArgs mfargs() const
    {
        MfArgs a;
        a.n = n;
        a.ku = nup;
        a.kd = ndown;
        a.na = na;
        a.nb = nb;
        a.N = N;
        a.unA = d_unA.p;
        a.unB = d_unB.p;
        a.rkA = d_rkA.p;
        a.rkB = d_rkB.p;
        a.h = d_h.p;
        a.g = d_g.p;
        a.diag = d_diag.p;
        a.FA = d_FA.p;
        a.FB = d_FB.p;
        return a;
    }
    size_t mf_lds() const { return (size_t)(nb + 2 * n * n) * sizeof(double) + ((size_t)2 << n); }

    DevSector sec() const
    {
        DevSector s;
        s.mode = mode;
        s.n = n;
        s.nel = nel;
        s.nb = nb;
        s.maskB = (n > 0) ? ((1u << n) - 1) : 0;
        s.states = d_states.p;
        s.rkA = d_rkA.p;
        s.rkB = d_rkB.p;
        s.binom = d_binom.p;
        return s;
    }

    SegArgs segargs(const Seg &sg, bool reverse) const
    {
        const TileSet &t = *tilesets[sg.tileset];
        SegArgs a;
        a.n = n;
        a.nb = nb;
        a.maskB = (1u << n) - 1;
        a.unA = d_unA.p;
        a.unB = d_unB.p;
        a.rkA = d_rkA.p;
        a.rkB = d_rkB.p;
        a.tileStart = t.tileStart.p;
        a.rows = t.rows.p;
        a.localRow = t.localRow.p;
        a.ops = d_ops.p;
        a.k0 = sg.k0;
        a.k1 = sg.k1;
        a.maxRows = t.maxRows;
        a.reverse = reverse ? 1 : 0;
        return a;
    }

    int orb_grid(const Seg &sg) const
    {
        return (int)std::max<long long>(1, std::min<long long>((sg.orbits + kBlock - 1) / kBlock, 2048));
    }

    OrbArgs orbargs(const Seg &sg, bool reverse, const uint32_t *progBase = nullptr) const
    {
        const TileSet &A = *tilesets[sg.orbA], &B = *tilesets[sg.orbB];
        OrbArgs a;
        a.n = n;
        a.nb = nb;
        a.maskB = (1u << n) - 1;
        a.unA = d_unA.p;
        a.unB = d_unB.p;
        a.aStart = A.tileStart.p;
        a.aRows = A.rows.p;
        a.bStart = B.tileStart.p;
        a.bRows = B.rows.p;
        a.aDesc = A.desc.p;
        a.bDesc = B.desc.p;
        a.nAc = A.ntiles;
        a.nBc = B.ntiles;
        a.ops = d_ops.p;
        a.k0 = sg.k0;
        a.k1 = sg.k1;
        a.reverse = reverse ? 1 : 0;
        a.prog = (progBase ? progBase : d_prog.p) + (sg.progOff >= 0 ? sg.progOff : 0);
        a.uaLocal = sg.ua;
        a.ubLocal = sg.ub;
        const uint32_t full = nq >= 32 ? 0xffffffffu : ((1u << nq) - 1);
        a.spect = full & ~((sg.ua << n) | sg.ub);
        return a;
    }

    size_t seg_lds(const Seg &sg, int nv) const
    {
        return (size_t)nv * tilesets[sg.tileset]->maxRows * nb * sizeof(double);
    }

    // Tiles for the alpha-bit set U (S_z sector).
    // Classes of strings of one spin agreeing outside the bits U (alpha when
    // beta == false): the tiles of an LDS segment and the orbit classes.
    int tileset_for(uint32_t U, std::string &err, bool beta = false)
    {
        const uint32_t key = U | (beta ? 0x80000000u : 0u);
        for (size_t i = 0; i < tilesets.size(); i++)
            if (tilesets[i]->U == key) return (int)i;
        auto t = std::make_unique<TileSet>();
        t->U = key;
        const std::vector<uint32_t> &un = beta ? unB : unA;
        const int cnt = beta ? nb : na;
        std::map<uint32_t, std::vector<int>> groups;  // spectator pattern -> ranks
        for (int ra = 0; ra < cnt; ra++) groups[un[ra] & ~U].push_back(ra);
        std::vector<int32_t> start, rows;
        std::vector<int8_t> local(cnt, 0);
        int maxRows = 0;
        for (auto &kv : groups) {
            start.push_back((int32_t)rows.size());
            int li = 0;
            for (int ra : kv.second) {  // ascending alpha rank within a tile
                rows.push_back(ra);
                local[ra] = (int8_t)li++;
            }
            maxRows = std::max(maxRows, li);
        }
        start.push_back((int32_t)rows.size());
        t->ntiles = (int)groups.size();
        t->maxRows = maxRows;
        if (maxRows > kMaxTileRows) {
            err = "tile too large";
            return -1;
        }
        if (!t->tileStart.ensure(start.size(), err) || !t->rows.ensure(rows.size(), err) ||
            !t->localRow.ensure(local.size(), err))
            return -1;
        if (maxRows <= 2) {
            std::vector<int4> d(t->ntiles);
            for (int c = 0; c < t->ntiles; c++) {
                const int r0 = rows[start[c]], r1 = (start[c + 1] - start[c] > 1) ? rows[start[c] + 1] : -1;
                d[c] = make_int4(r0, r1, (int)un[r0], r1 >= 0 ? (int)un[r1] : (int)un[r0]);
            }
            if (!t->desc.ensure(d.size(), err) ||
                hipMemcpy(t->desc.p, d.data(), d.size() * sizeof(int4), hipMemcpyHostToDevice) != hipSuccess) {
                err = "class descriptor upload failed";
                return -1;
            }
        }
        if (hipMemcpy(t->tileStart.p, start.data(), start.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(t->rows.p, rows.data(), rows.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(t->localRow.p, local.data(), local.size(), hipMemcpyHostToDevice) != hipSuccess) {
            err = "tile upload failed";
            return -1;
        }
        tilesets.push_back(std::move(t));
        return (int)tilesets.size() - 1;
    }

    void fill_cs(const double *angles, int nvec)
    {
        h_cs.resize((size_t)nvec * M);
        for (size_t i = 0; i < (size_t)nvec * M; i++) {
            double s, c;
            // the reference's mysincos (src/linalg.h:31-35)
            __builtin_sincos(angles[i], &s, &c);
            h_cs[i] = make_double2(c, s);
        }
    }

    bool upload_cs(const double *angles, int nvec, DBuf<double2> &fwd, DBuf<double2> &rev, std::string &err)
    {
        fill_cs(angles, nvec);
        const size_t cnt = (size_t)nvec * M;
        if (!fwd.ensure(cnt, err) || !rev.ensure(cnt, err)) return false;
        // sincos(-x) = (-sin x, cos x): the reverse sweep's negated angles
        h_csneg.resize(cnt);
        for (size_t i = 0; i < cnt; i++) h_csneg[i] = make_double2(h_cs[i].x, -h_cs[i].y);
        HCHK(hipMemcpyAsync(fwd.p, h_cs.data(), cnt * sizeof(double2), hipMemcpyHostToDevice, stream));
        HCHK(hipMemcpyAsync(rev.p, h_csneg.data(), cnt * sizeof(double2), hipMemcpyHostToDevice, stream));
        HCHK(hipStreamSynchronize(stream));  // the host tables are reused by the next call
        return true;
    }

    // Angle tables of the single-vector sequences (d_cs, d_csneg): with
    // complex amplitudes one table per plane, both the same.
    bool upload_cs_single(const double *angles, std::string &err)
    {
        if (!cplx) return upload_cs(angles, 1, d_cs, d_csneg, err);
        std::vector<double> two(2 * (size_t)M);
        std::copy(angles, angles + M, two.begin());
        std::copy(angles, angles + M, two.begin() + M);
        return upload_cs(two.data(), 2, d_cs, d_csneg, err);
    }

    // Results of the single-vector E + gradient sequence: d_out holds
    // planes() rows of M per-operator derivatives, then planes() energies;
    // the planes add up.
    bool collect_single(double *deriv, double *E, std::string &err)
    {
        const int pl = planes();
        std::vector<double> out((size_t)pl * (M + 1));
        HCHK(hipMemcpyAsync(out.data(), d_out.p, 8 * out.size(), hipMemcpyDeviceToHost, stream));
        HCHK(hipStreamSynchronize(stream));
        collect_marks();
        for (int k = 0; k < M; k++) deriv[k] = pl == 2 ? out[k] + out[M + k] : out[k];
        if (E) *E = pl == 2 ? out[2 * M] + out[2 * M + 1] : out[M];
        return true;
    }

    // source rows per phase of the vector-major split sigma (kOptSigma1Rows),
    // reduced to what the device's LDS holds; 0 = the two-row kernel.
    // Measured, one vector: 853,776 determinants 0.80 ms (two rows) ->
    // 0.49 ms (8 rows); 11,778,624: 16.1 -> 10.8 ms (4 rows, the most that fit)
    int sig1Rows = 8;
    int sig1_rows() const
    {
        int rb = sig1Rows;
        while (rb > 2 && (size_t)rb * (kSig1GS + nb) * sizeof(double) > ldsDevice) rb -= 2;
        return rb;
    }

    // An error for the operations that work on real amplitudes only.
    bool real_only(const char *what, std::string &err) const
    {
        if (!cplx) return true;
        err = std::string(what) + " is not available for complex amplitudes";
        return false;
    }

    // Forward sweep of nvec vectors with per-vector angle tables.
    // k_orb2s grid of a vector-major two-group segment (<= slots)
    int s_grid(const Seg &sg) const
    {
        return (int)std::max<long long>(1, std::min<long long>((sg.orbits + kBlock - 1) / kBlock, slots));
    }

    bool sweep_forward(double *v, int nvec, const double2 *cs, std::string &err)
    {
        if (use_wv()) return sweep_forward_wv(v, nvec, cs, err);
        const int mk = mark_begin(1);
        if (use_s()) {
            for (const Seg &sg : laneS.segs) {
                if (sg.two) {
                    Orb2Args oa = orb2args(laneS, sg, false);
                    oa.list = oa.llist = nullptr;
                    hipLaunchKernelGGL((k_orb2s<false>), dim3(s_grid(sg), nvec), dim3(kBlock), 0, stream, oa, v,
                                       (double *)nullptr, N, cs, M, (double *)nullptr, slots);
                } else if (sg.orbA >= 0) {
                    hipLaunchKernelGGL((k_orb<false>), dim3(std::min(orb_grid(sg), slots), nvec), dim3(kBlock), 0,
                                       stream, orbargs(sg, false, laneS.lprog.p), v, nullptr, N, cs, (long long)M,
                                       nullptr, 0, 0);
                } else {
                    for (int k = sg.k0; k < sg.k1; k++)
                        hipLaunchKernelGGL((k_op_generic<false>), dim3(Ggen, nvec), dim3(kBlock), 0, stream, v,
                                           nullptr, N, N, sec(), d_ops.p, k, cs, (long long)M, nullptr, 0, 0);
                }
                LCHK();
            }
            mark_end(mk);
            return true;
        }
        for (const Seg &sg : segs) {
            if (sg.diag) {
                hipLaunchKernelGGL((k_phase<false>), dim3(Ggen, nvec / 2), dim3(kBlock), 0, stream, v, nullptr, N,
                                   N, d_states.p, d_ops.p, sg.k0, sg.k1, cs, (long long)M, nullptr, 0, 0);
            } else if (sg.orbA >= 0 && orbLds) {
                hipLaunchKernelGGL((k_orbl<false>), dim3(tilesets[sg.orbA]->ntiles, nvec), dim3(kBlock),
                                   (size_t)2 * nb * sizeof(double), stream, orbargs(sg, false), v, nullptr, N, cs,
                                   (long long)M, nullptr, 0, 0);
            } else if (sg.orbA >= 0) {
                hipLaunchKernelGGL((k_orb<false>), dim3(orb_grid(sg), nvec), dim3(kBlock), 0, stream,
                                   orbargs(sg, false), v, nullptr, N, cs, (long long)M, nullptr, 0, 0);
            } else if (sg.tileset >= 0) {
                const TileSet &t = *tilesets[sg.tileset];
                hipLaunchKernelGGL((k_seg<1, false>), dim3(t.ntiles, nvec), dim3(kBlock), seg_lds(sg, 1), stream,
                                   segargs(sg, false), v, nullptr, N, cs, (long long)M, nullptr, 0, 0);
            } else {
                hipLaunchKernelGGL((k_op_generic<false>), dim3(Ggen, nvec), dim3(kBlock), 0, stream, v, nullptr, N,
                                   N, sec(), d_ops.p, sg.k0, cs, (long long)M, nullptr, 0, 0);
            }
            LCHK();
        }
        mark_end(mk);
        return true;
    }

    // Adjoint sweep of (phi, lambda) pairs with per-operator brackets.
    bool sweep_reverse(double *phi, double *lam, int nvec, const double2 *csneg, double *partial, std::string &err)
    {
        const long long pstride = (long long)M * slots;
        const long long total = pstride * nvec;
        hipLaunchKernelGGL(k_fill, dim3((unsigned)std::min<long long>((total + kBlock - 1) / kBlock, 4096)),
                           dim3(kBlock), 0, stream, partial, total, 0.0);
        LCHK();
        if (use_wv()) return sweep_reverse_wv(phi, lam, nvec, csneg, partial, err);
        const int mk = mark_begin(2);
        if (use_s()) {
            for (int si = (int)laneS.segs.size() - 1; si >= 0; si--) {
                const Seg &sg = laneS.segs[si];
                if (sg.two) {
                    Orb2Args oa = orb2args(laneS, sg, true);
                    oa.list = oa.llist = nullptr;
                    hipLaunchKernelGGL((k_orb2s<true>), dim3(s_grid(sg), nvec), dim3(kBlock), 0, stream, oa, phi,
                                       lam, N, csneg, M, partial, slots);
                } else if (sg.orbA >= 0) {
                    hipLaunchKernelGGL((k_orb<true>), dim3(std::min(orb_grid(sg), slots), nvec), dim3(kBlock), 0,
                                       stream, orbargs(sg, true, laneS.lprog.p), phi, lam, N, csneg, (long long)M,
                                       partial, slots, pstride);
                } else {
                    for (int k = sg.k1 - 1; k >= sg.k0; k--)
                        hipLaunchKernelGGL((k_op_generic<true>), dim3(Ggen, nvec), dim3(kBlock), 0, stream, phi, lam,
                                           N, N, sec(), d_ops.p, k, csneg, (long long)M, partial, slots, pstride);
                }
                LCHK();
            }
            mark_end(mk);
            return true;
        }
        for (int si = (int)segs.size() - 1; si >= 0; si--) {
            const Seg &sg = segs[si];
            if (sg.diag) {
                hipLaunchKernelGGL((k_phase<true>), dim3(Ggen, nvec / 2), dim3(kBlock), 0, stream, phi, lam, N, N,
                                   d_states.p, d_ops.p, sg.k0, sg.k1, csneg, (long long)M, partial, slots, pstride);
            } else if (sg.orbA >= 0 && orbLds) {
                hipLaunchKernelGGL((k_orbl<true>), dim3(tilesets[sg.orbA]->ntiles, nvec), dim3(kBlock),
                                   (size_t)4 * nb * sizeof(double), stream, orbargs(sg, true), phi, lam, N, csneg,
                                   (long long)M, partial, slots, pstride);
            } else if (sg.orbA >= 0) {
                hipLaunchKernelGGL((k_orb<true>), dim3(orb_grid(sg), nvec), dim3(kBlock), 0, stream,
                                   orbargs(sg, true), phi, lam, N, csneg, (long long)M, partial, slots, pstride);
            } else if (sg.tileset >= 0) {
                const TileSet &t = *tilesets[sg.tileset];
                hipLaunchKernelGGL((k_seg<2, true>), dim3(t.ntiles, nvec), dim3(kBlock), seg_lds(sg, 2), stream,
                                   segargs(sg, true), phi, lam, N, csneg, (long long)M, partial, slots, pstride);
            } else {
                hipLaunchKernelGGL((k_op_generic<true>), dim3(Ggen, nvec), dim3(kBlock), 0, stream, phi, lam, N, N,
                                   sec(), d_ops.p, sg.k0, csneg, (long long)M, partial, slots, pstride);
            }
            LCHK();
        }
        mark_end(mk);
        return true;
    }

    // sin/cos tables of a batch in lane layout, padding lanes (1, 0)
    bool upload_cs_lanes(const double *angles, int nvec, std::string &err)
    {
        fill_cs(angles, nvec);
        const int nch = (nvec + kLanes - 1) / kLanes;
        const size_t cnt = (size_t)nch * M * kLanes;
        if (!bl_cs.ensure(cnt, err) || !bl_csneg.ensure(cnt, err)) return false;
        h_csl.assign(cnt, make_double2(1.0, 0.0));
        h_cslneg.assign(cnt, make_double2(1.0, 0.0));
        for (int v = 0; v < nvec; v++) {
            const int c = v / kLanes, l = v % kLanes;
            for (int k = 0; k < M; k++) {
                const double2 t = h_cs[(size_t)v * M + k];
                const size_t o = ((size_t)c * M + k) * kLanes + l;
                h_csl[o] = t;
                h_cslneg[o] = make_double2(t.x, -t.y);
            }
        }
        HCHK(hipMemcpyAsync(bl_cs.p, h_csl.data(), cnt * sizeof(double2), hipMemcpyHostToDevice, stream));
        HCHK(hipMemcpyAsync(bl_csneg.p, h_cslneg.data(), cnt * sizeof(double2), hipMemcpyHostToDevice, stream));
        HCHK(hipStreamSynchronize(stream));
        return true;
    }

    // state s of the stored forward states (checkpoint mode)
    double *ck_state(int s, int nch) { return bl_ck.p + (size_t)s * nch * N * kLanes; }
    // forward two-group launch grid (smaller with a reachable-orbit list)
    static int fgrid(const Seg &sg) { return sg.supOff >= 0 ? sg.gridSup : sg.grid; }

    // Whether the first forward segment reads the initial state directly (one
    // plain vector broadcast to every lane) instead of a lane-layout state 0:
    // checkpoint mode with a two-group first segment.
    bool first_reads_start() const
    {
        if (winOn) return ckOn;  // window pass 0 reads the start vector
        return ckOn && !laneF.segs.empty() && laneF.segs[0].two && !laneF.segs[0].three;
    }

    // Forward sweep of lane-layout chunks: in place on v, or (checkpoint
    // mode) segment s from state s to state s + 1, v being state 0 (with
    // first_reads_start(), segment 0 reads d_start instead and state 0 is
    // never written).  With timing on, every launch is bracketed by its own
    // events.
    bool sweep_forward_lanes(double *v, int nch, std::string &err)
    {
        if (winOn) return sweep_forward_win(v, nch, err);
        const long long cstride = N * kLanes;
        for (size_t si = 0; si < laneF.segs.size(); si++) {
            const Seg &sg = laneF.segs[si];
            const double *in = ckOn ? ck_state((int)si, nch) : v;
            double *out = ckOn ? ck_state((int)si + 1, nch) : v;
            if (si == 0 && first_reads_start()) {
                const int mk = mark_begin(1);
                hipLaunchKernelGGL((k_orb2v<false, true>), dim3(fgrid(sg), nch), dim3(kBlock), 0, stream,
                                   orb2args(laneF, sg, false), d_start.p, out, nullptr, cstride, bl_cs.p, M, nullptr,
                                   lslotsF);
                mark_end(mk);
                LCHK();
            } else if (sg.three) {
                const int mk = mark_begin(1);
                hipLaunchKernelGGL((k_orb3v<false>), dim3(sg.grid, nch), dim3(kBlock), 0, stream, orb3args(laneF, sg),
                                   in, out, nullptr, cstride, bl_cs.p, M, nullptr, lslotsF);
                mark_end(mk);
                LCHK();
            } else if (sg.two) {
                const int mk = mark_begin(1);
                hipLaunchKernelGGL((k_orb2v<false>), dim3(fgrid(sg), nch), dim3(kBlock), 0, stream,
                                   orb2args(laneF, sg, false), in, out, nullptr, cstride, bl_cs.p, M, nullptr,
                                   lslotsF);
                mark_end(mk);
                LCHK();
            } else if (sg.orbA >= 0) {
                const int mk = mark_begin(1);
                hipLaunchKernelGGL((k_orbv<false>), dim3(sg.grid, nch), dim3(kBlock), 0, stream,
                                   orbargs(sg, false, laneF.lprog.p), in, out, nullptr, cstride, bl_cs.p, M,
                                   nullptr, lslotsF);
                mark_end(mk);
                LCHK();
            } else {
                for (int k = sg.k0; k < sg.k1; k++) {
                    const int mk = mark_begin(1);
                    hipLaunchKernelGGL((k_opv<false>), dim3(sg.grid, nch), dim3(kBlock), 0, stream, v, nullptr,
                                       cstride, N, sec(), d_ops.p, k, bl_cs.p, M, nullptr, lslotsF);
                    mark_end(mk);
                    LCHK();
                }
            }
        }
        return true;
    }

    // Adjoint sweep of lane-layout (phi, lambda) chunks; every operator's
    // partial rows are written by exactly one launch (rows 0 .. its grid - 1).
    // Checkpoint mode: segment s reads phi from stored state s + 1 and does
    // not write it back (only lambda is updated in place).
    // energy rows of the fused first adjoint pass (register window tiles)
    bool fused_energy() const { return winOn && winReg && !wpass.empty(); }
    int win_erows() const { return wpass.empty() ? 0 : wpass.back().gridA; }
    // the energy rows' groups (as the kernel forms them)
    int win_egroups() const { return 1 << std::min(4, __builtin_ctz((unsigned)std::max(1, win_erows()))); }
    bool sweep_reverse_lanes(double *phi, double *lam, int nch, std::string &err, double *epart = nullptr)
    {
        if (winOn) return sweep_reverse_win(phi, lam, nch, err, epart);
        const long long cstride = N * kLanes;
        for (int si = (int)lsegs.size() - 1; si >= 0; si--) {
            const Seg &sg = lsegs[si];
            const double *in = ckOn ? ck_state(si + 1, nch) : phi;
            double *out = ckOn ? nullptr : phi;
            if (sg.three) {
                const int mk = mark_begin(2);
                hipLaunchKernelGGL((k_orb3v<true>), dim3(sg.grid, nch), dim3(kBlock), 0, stream, orb3args(laneA, sg),
                                   in, out, lam, cstride, bl_csneg.p, M, bl_partial.p, lslots);
                mark_end(mk);
                LCHK();
            } else if (sg.two) {
                const int mk = mark_begin(2);
                const Orb2Args oa = orb2args(laneA, sg, true);
                if (oa.list)
                    hipLaunchKernelGGL((k_orb2v<true, false, true>), dim3(sg.grid, nch), dim3(kBlock), 0, stream, oa,
                                       in, out, lam, cstride, bl_csneg.p, M, bl_partial.p, lslots);
                else
                    hipLaunchKernelGGL((k_orb2v<true>), dim3(sg.grid, nch), dim3(kBlock), 0, stream, oa, in, out,
                                       lam, cstride, bl_csneg.p, M, bl_partial.p, lslots);
                mark_end(mk);
                LCHK();
            } else if (sg.orbA >= 0) {
                const int mk = mark_begin(2);
                hipLaunchKernelGGL((k_orbv<true>), dim3(sg.grid, nch), dim3(kBlock), 0, stream,
                                   orbargs(sg, true, laneA.lprog.p), in, out, lam, cstride, bl_csneg.p, M,
                                   bl_partial.p, lslots);
                mark_end(mk);
                LCHK();
            } else {
                for (int k = sg.k1 - 1; k >= sg.k0; k--) {
                    const int mk = mark_begin(2);
                    hipLaunchKernelGGL((k_opv<true>), dim3(sg.grid, nch), dim3(kBlock), 0, stream, phi, lam, cstride,
                                       N, sec(), d_ops.p, k, bl_csneg.p, M, bl_partial.p, lslots);
                    mark_end(mk);
                    LCHK();
                }
            }
        }
        return true;
    }

    // The same-spin DGEMMs of the lane-layout sigma with rocBLAS solution sol
    // (0: the library's default).  which 0, beta part: per alpha row,
    // Y_a (nb x 64) += V_bb X_a (column-major view: Y_a^T += X_a^T V_bb^T);
    // which 1, alpha part: per chunk, Y (na x nb*64) += V_aa X.
    // overwrite: y = V x (beta 0) instead of y += V x
    rocblas_status sigma_gemm(int which, const double *x, double *y, int nch, int sol, bool list = false,
                              rocblas_int *sols = nullptr, rocblas_int *nsol = nullptr, bool overwrite = false)
    {
        const double one = 1.0, zero = 0.0;
        const double *beta = overwrite ? &zero : &one;
        const rocblas_datatype f = rocblas_datatype_f64_r;
        const rocblas_gemm_algo algo = (sol != 0 || list) ? rocblas_gemm_algo_solution_index
                                                          : rocblas_gemm_algo_standard;
        const int m = which ? nb * kLanes : kLanes, nk = which ? na : nb;
        const int ld = which ? nb * kLanes : kLanes;
        const rocblas_stride st = which ? (rocblas_stride)N * kLanes : (rocblas_stride)nb * kLanes;
        const double *V = which ? d_Vaa.p : d_Vbb.p;
        const int batch = which ? nch : na * nch;
        if (list)
            return rocblas_gemm_strided_batched_ex_get_solutions(
                blas, rocblas_operation_none, rocblas_operation_none, m, nk, nk, &one, x, f, ld, st, V, f, nk, 0, &one,
                y, f, ld, st, y, f, ld, st, batch, f, algo, 0, sols, nsol);
        return rocblas_gemm_strided_batched_ex(blas, rocblas_operation_none, rocblas_operation_none, m, nk, nk, &one,
                                               x, f, ld, st, V, f, nk, 0, beta, y, f, ld, st, y, f, ld, st, batch, f,
                                               algo, sol, 0);
    }

    // Lane-layout sigma for nch chunks and the per-vector energies
    // out[v] = <x_v|y_v> (matrix-free storage, n <= 12).
    // epart: 512 x nch x 64 energy partials (the batch's bl_epart when null)
    // dot false: no energies (the batch takes them from the first adjoint
    // window pass)
    // z (exchange-symmetric batches with the grouped sigma): the alpha-alpha
    // DGEMM writes z = V_aa x first, and the grouped sigma adds z[a][b] +
    // z[b][a] (both same-spin parts) to its sums; no beta-beta DGEMM
    bool ham_apply_lanes(const double *x, double *y, int nch, int nvec, double *out, std::string &err,
                         double *epart = nullptr, bool dot = true, double *z = nullptr)
    {
        if (!epart) epart = bl_epart.p;
        if (z && !(sigg && siggOpt && na == nb)) z = nullptr;
        if (z) {
            const int mk3 = mark_begin(3);
            rocblas_status st = sigma_gemm(1, x, z, nch, gemmSol[1], false, nullptr, nullptr, true);
            if (st != rocblas_status_success && gemmSol[1] != 0) {
                gemmSol[1] = 0;
                st = sigma_gemm(1, x, z, nch, 0, false, nullptr, nullptr, true);
            }
            if (st != rocblas_status_success) {
                err = "rocBLAS DGEMM (alpha part) failed";
                return false;
            }
            mark_end(mk3);
        }
        const int mk = mark_begin(0);
        SigvArgs s;
        s.m = mfargs();
        s.ts = d_ts.p;
        s.src = d_sigsrc.p;
        s.ntp = ntp16;
        s.x = x;
        s.y = y;
        s.z = z;
        s.cstride = N * kLanes;
        const unsigned gx = (unsigned)((na + 7) / 8 * 32);
        if (sigg && siggOpt) {
            hipLaunchKernelGGL(k_sigg_sparse<6>, dim3(gx, nch), dim3(kSigWThreads), kSigWLds, stream, s);
            if (z) {
                // the targets below the diagonal from their mirror images
                LCHK();
                hipLaunchKernelGGL(k_mirror_lanes, dim3((unsigned)na, nch), dim3(kBlock), 0, stream, y, nb,
                                   s.cstride);
            }
        } else if (sigh && sighOpt) {
            SighArgs h;
            h.m = s.m;
            h.ts = d_tsh.p;
            h.hl = d_hl.p;
            h.nA = sighNA;
            h.nB = sighNB;
            h.TA = sighTA;
            h.T = sighT;
            h.x = x;
            h.y = y;
            h.cstride = s.cstride;
            hipLaunchKernelGGL(sigh_kernel(sighS, sighX), dim3(gx, nch), dim3(kSigWThreads), kSigHLds, stream, h);
        } else {
            hipLaunchKernelGGL(sigw_kernel(ntp16 / 4), dim3(gx, nch), dim3(kSigWThreads), kSigWLds, stream, s);
        }
        mark_end(mk);
        LCHK();
        if (!z) {
            const int mk3 = mark_begin(3);
            for (int which = 0; which < 2; which++) {
                rocblas_status st = sigma_gemm(which, x, y, nch, gemmSol[which]);
                if (st != rocblas_status_success && gemmSol[which] != 0) {
                    // a recorded solution this library does not offer: the default
                    gemmSol[which] = 0;
                    st = sigma_gemm(which, x, y, nch, 0);
                }
                if (st != rocblas_status_success) {
                    err = which ? "rocBLAS DGEMM (alpha part) failed" : "rocBLAS DGEMM (beta part) failed";
                    return false;
                }
            }
            mark_end(mk3);
        }
        if (!dot) return true;
        const int G = 512;
        hipLaunchKernelGGL(k_dot_lanes, dim3(G, nch), dim3(kBlock), 0, stream, x, y, N, epart);
        LCHK();
        hipLaunchKernelGGL(k_reduce_lanes, dim3(nch), dim3(kBlock), 0, stream, epart, 1, G, nullptr, nvec, out, 1.0);
        LCHK();
        return true;
    }

    // y = M^T x for nvec vectors (stride N); energy partials epart[v][eslots].
    // half: the vectors are exchange-symmetric (x[a][b] == x[b][a]); the
    // vector-major split sigma then computes the targets b >= a and copies
    // the rest from their mirror images
    bool ham_apply(const double *x, double *y, int nvec, double *epart, std::string &err, bool half = false)
    {
        const int es = eslots();
        if (hkind == 2) {
            int v = 0;
            while (v < nvec) {
                const int left = nvec - v;
                const double *xv = x + (size_t)v * N;
                double *yv = y + (size_t)v * N;
                double *ev = epart + (size_t)v * es;
                const int mk = mark_begin(0);
                if (left >= 8) {
                    hipLaunchKernelGGL(k_ell<8>, dim3(es, left / 8), dim3(kBlock), 0, stream, d_col.p, d_val.p, W, N,
                                       xv, yv, N, ev, es);
                    v += (left / 8) * 8;
                } else if (left >= 4) {
                    hipLaunchKernelGGL(k_ell<4>, dim3(es, 1), dim3(kBlock), 0, stream, d_col.p, d_val.p, W, N, xv,
                                       yv, N, ev, es);
                    v += 4;
                } else if (left >= 2) {
                    hipLaunchKernelGGL(k_ell<2>, dim3(es, 1), dim3(kBlock), 0, stream, d_col.p, d_val.p, W, N, xv,
                                       yv, N, ev, es);
                    v += 2;
                } else {
                    hipLaunchKernelGGL(k_ell<1>, dim3(es, 1), dim3(kBlock), 0, stream, d_col.p, d_val.p, W, N, xv,
                                       yv, N, ev, es);
                    v += 1;
                }
                mark_end(mk);
                LCHK();
            }
            return true;
        }
        if (hkind == 3 && sig1) {
            // split sigma: sparse part, then the same-spin parts as DGEMMs
            // over whole vectors, then the energy partials
            int mk = mark_begin(0);
            Sig1Args s;
            s.m = mfargs();
            s.ts = d_ts1.p;
            s.ntp = ntp;
            s.x = x;
            s.y = y;
            s.vstride = N;
            const int rb = sig1_rows();
            s.half = half && rb > 0 && na == nb ? 1 : 0;
            const size_t lds = (size_t)rb * (kSig1GS + nb) * sizeof(double);
            if (rb == 8)
                hipLaunchKernelGGL(k_sig1r<8>, dim3(na, nvec), dim3(kSig1Threads), lds, stream, s);
            else if (rb == 6)
                hipLaunchKernelGGL(k_sig1r<6>, dim3(na, nvec), dim3(kSig1Threads), lds, stream, s);
            else if (rb == 4)
                hipLaunchKernelGGL(k_sig1r<4>, dim3(na, nvec), dim3(kSig1Threads), lds, stream, s);
            else if (rb == 2)
                hipLaunchKernelGGL(k_sig1r<2>, dim3(na, nvec), dim3(kSig1Threads), lds, stream, s);
            else
                hipLaunchKernelGGL(k_sig1_sparse, dim3(na, nvec), dim3(kSig1Threads), kSig1Lds, stream, s);
            if (s.half) {
                LCHK();
                const unsigned tb = (unsigned)((nb + 31) / 32);
                hipLaunchKernelGGL(k_mirror_rows, dim3(tb, tb, nvec), dim3(kBlock), 0, stream, y, nb, (long long)N);
            }
            mark_end(mk);
            LCHK();
            mk = mark_begin(3);
            const double one = 1.0;
            // alpha part: Y_v (na x nb) += V_aa X_v   (column-major: Y^T += X^T V_aa^T)
            if (rocblas_dgemm_strided_batched(blas, rocblas_operation_none, rocblas_operation_none, nb, na, na, &one,
                                              x, nb, (rocblas_stride)N, d_Vaa.p, na, 0, &one, y, nb,
                                              (rocblas_stride)N, nvec) != rocblas_status_success) {
                err = "rocblas_dgemm_strided_batched (alpha part) failed";
                return false;
            }
            // beta part: Y_v += X_v V_bb^T   (column-major: Y^T += V_bb X^T)
            if (rocblas_dgemm_strided_batched(blas, rocblas_operation_transpose, rocblas_operation_none, nb, na, nb,
                                              &one, d_Vbb.p, nb, 0, x, nb, (rocblas_stride)N, &one, y, nb,
                                              (rocblas_stride)N, nvec) != rocblas_status_success) {
                err = "rocblas_dgemm_strided_batched (beta part) failed";
                return false;
            }
            mark_end(mk);
            hipLaunchKernelGGL(k_dot_rows, dim3(eslots(), nvec), dim3(kBlock), 0, stream, x, y, N, N, epart);
            LCHK();
            return true;
        }
        if (hkind == 3) {
            const int mk = mark_begin(0);
            hipLaunchKernelGGL(k_mf_sigma, dim3(na, nvec), dim3(kBlock), mf_lds(), stream, mfargs(), x, y, N, epart);
            mark_end(mk);
            LCHK();
            return true;
        }
        if (hkind == 1) {
            const int mk = mark_begin(0);
            hipLaunchKernelGGL(k_rows, dim3(es, nvec), dim3(kBlock), 0, stream, d_rowptr.p, d_col.p, d_val.p, N, x, y,
                               N, epart, es);
            mark_end(mk);
            LCHK();
            return true;
        }
        err = "no Hamiltonian on the device";
        return false;
    }
};

Engine::Engine() : d_(new EngineImpl) {}

Engine::~Engine()
{
    if (d_) d_->drop_graphs();
    if (d_ && d_->blas) (void)rocblas_destroy_handle(d_->blas);
    if (d_ && d_->stream) (void)hipStreamDestroy(d_->stream);
}

long long Engine::dim() const { return d_->N; }
int Engine::mode() const { return d_->mode; }
const std::vector<uint32_t> &Engine::states() const { return d_->states; }
long long Engine::ham_nnz() const { return d_->nnz; }
int Engine::ham_kind() const { return d_->hkind; }
int Engine::num_segments() const { return (int)d_->segs.size(); }
int Engine::num_generic_ops() const { return d_->ngeneric; }

long long Engine::rank(uint32_t b) const
{
    const auto &st = d_->states;
    auto it = std::lower_bound(st.begin(), st.end(), b);
    if (it == st.end() || *it != b) return -1;
    return (long long)(it - st.begin());
}

bool Engine::init(int nq, int mode, int nel, int nup, int ndown, std::string &err)
{
    EngineImpl &e = *d_;
    if (nq < 1 || nq > 30) {
        err = "number of qubits must be 1..30";
        return false;
    }
    int dev = 0;
    HCHK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HCHK(hipGetDeviceProperties(&prop, dev));
    e.ldsMax = prop.sharedMemPerBlock;
    e.ldsDevice = prop.sharedMemPerBlock;
    e.numCU = std::max(1, prop.multiProcessorCount);
    if (const char *b = getenv("AE_LDS_BUDGET")) e.ldsMax = std::min<size_t>(e.ldsMax, (size_t)atol(b));
    HCHK(hipStreamCreateWithFlags(&e.stream, hipStreamNonBlocking));
    if (const char *g = getenv("AE_GRAPH")) e.useGraph = std::string(g) != "0";
    if (const char *sw = getenv("AE_SWEEP")) {
        const std::string v(sw);
        e.sweepKind = v == "tile" ? kSweepTile : v == "orblds" ? kSweepOrbitLds : v == "orbit2" ? kSweepOrbit2 : kSweepOrbit;
    }
    if (const char *l = getenv("AE_LANES")) e.laneMin = std::max(0, atoi(l));
    if (const char *r = getenv("AE_REACH")) e.reachOpt = atoi(r) ? 1 : 0;
    e.nq = nq;
    e.mode = mode;
    e.nel = nel;
    e.nup = nup;
    e.ndown = ndown;
    e.n = nq / 2;
    if (mode == kSz && (nq % 2 != 0)) {
        err = "S_z sector needs an even qubit count";
        return false;
    }
    // sector bit strings, ascending (src/operatorpool.cpp:207-242,
    // src/operatorpool.h:49-77)
    e.states.clear();
    if (mode == kSz) {
        const int n = e.n;
        e.rkA.assign(1u << n, -1);
        e.rkB.assign(1u << n, -1);
        e.unA.clear();
        e.unB.clear();
        for (uint32_t s = 0; s < (1u << n); s++) {
            if (__builtin_popcount(s) == nup) {
                e.rkA[s] = (int32_t)e.unA.size();
                e.unA.push_back(s);
            }
            if (__builtin_popcount(s) == ndown) {
                e.rkB[s] = (int32_t)e.unB.size();
                e.unB.push_back(s);
            }
        }
        e.na = (int)e.unA.size();
        e.nb = (int)e.unB.size();
        e.N = (long long)e.na * e.nb;
        e.states.resize(e.N);
        for (long long i = 0; i < e.N; i++)
            e.states[i] = (e.unA[i / e.nb] << n) | e.unB[i % e.nb];
        if (!e.d_unA.ensure(e.na, err) || !e.d_unB.ensure(e.nb, err) || !e.d_rkA.ensure(1u << n, err) ||
            !e.d_rkB.ensure(1u << n, err))
            return false;
        HCHK(hipMemcpy(e.d_unA.p, e.unA.data(), 4 * e.na, hipMemcpyHostToDevice));
        HCHK(hipMemcpy(e.d_unB.p, e.unB.data(), 4 * e.nb, hipMemcpyHostToDevice));
        HCHK(hipMemcpy(e.d_rkA.p, e.rkA.data(), 4 * (1u << n), hipMemcpyHostToDevice));
        HCHK(hipMemcpy(e.d_rkB.p, e.rkB.data(), 4 * (1u << n), hipMemcpyHostToDevice));
    } else {
        const uint32_t full = 1u << nq;
        if (mode == kNumber) {
            e.states.reserve((size_t)binom_ll(nq, nel));
            for (uint32_t s = 0; s < full; s++)
                if (__builtin_popcount(s) == nel) e.states.push_back(s);
        } else {
            e.states.resize(full);
            std::iota(e.states.begin(), e.states.end(), 0u);
        }
        e.N = (long long)e.states.size();
        std::vector<uint32_t> bn(33 * 33, 0);
        for (int p = 0; p < 33; p++)
            for (int t = 0; t < 33; t++) bn[p * 33 + t] = (uint32_t)std::min<long long>(binom_ll(p, t), 0xffffffffll);
        if (!e.d_binom.ensure(bn.size(), err)) return false;
        HCHK(hipMemcpy(e.d_binom.p, bn.data(), 4 * bn.size(), hipMemcpyHostToDevice));
    }
    if (!e.d_states.ensure(e.N, err)) return false;
    HCHK(hipMemcpy(e.d_states.p, e.states.data(), 4 * e.N, hipMemcpyHostToDevice));
    e.Ggen = (int)std::min<long long>((e.N + kBlock - 1) / kBlock, 1024);
    // vector-major sweeps: two TUPS units per pass (half the launches) where
    // a pass is long enough to outweigh the heavier thread-per-orbit kernel
    // (measured, one vector: 853,776 determinants 3.11 -> 2.49 ms per E+grad;
    // 63,504: 0.78 -> 0.83 ms)
    if (!getenv("AE_SWEEP")) e.sweepKind = (mode == kSz && e.N >= (1LL << 18)) ? kSweepOrbit2 : kSweepOrbit;
    if (!e.d_psi.ensure(e.N, err) || !e.d_phi.ensure(e.N, err) || !e.d_lam.ensure(e.N, err) ||
        !e.d_start.ensure(e.N, err))
        return false;
    return true;
}

bool Engine::set_operators(const std::vector<int8_t> &ops4, const std::vector<int> &order,
                           const std::vector<double> &scale, int nparams, std::string &err)
{
    EngineImpl &e = *d_;
    const int M = (int)(ops4.size() / 4);
    e.drop_graphs();
    if (&ops4 != &e.ops4) e.ops4 = ops4;
    e.M = M;
    e.P = nparams;
    e.order = order;
    e.scale = scale;
    e.hops.resize(M);
    for (int k = 0; k < M; k++)
        if (!make_devop(&ops4[4 * k], e.nq, e.hops[k], err, e.cplx)) {
            err = "operator " + std::to_string(k + 1) + ": " + err;
            return false;
        }
    if (!e.d_ops.ensure(M, err)) return false;
    HCHK(hipMemcpy(e.d_ops.p, e.hops.data(), sizeof(DevOp) * M, hipMemcpyHostToDevice));
    e.find_mirrors();

    // Fused segments (S_z sector): maximal runs of consecutive operators whose
    // union U of alpha bits keeps a tile of two vectors within the LDS budget.
    // Operator order is preserved, so the unitary is unchanged.
    e.segs.clear();
    e.ngeneric = 0;
    int maxTiles = 0;
    // number operators (complex amplitudes only): a run of up to
    // kPhaseMaxOps consecutive ones is one k_phase segment, and no other
    // segment takes one in
    auto isDiag = [&](int kk) { return e.hops[kk].cre == e.hops[kk].ann; };
    auto diagRun = [&](int kk) {
        int k1 = kk + 1;
        while (k1 < M && isDiag(k1) && k1 - kk < kPhaseMaxOps) k1++;
        Seg sg{kk, k1, -1};
        sg.diag = 1;
        e.segs.push_back(sg);
        return k1;
    };
    if (e.mode == kSz) {
        const int n = e.n;
        const uint32_t amask = (1u << n) - 1;
        auto rows_for = [&](int u) {
            long long r = 0;
            for (int j = 0; j <= u; j++) r = std::max(r, binom_ll(u, j));
            return r;
        };
        auto fits = [&](uint32_t U) {
            const int u = __builtin_popcount(U);
            const long long rows = rows_for(u);
            return rows <= kMaxTileRows && (size_t)(2 * rows * e.nb * sizeof(double)) <= e.ldsMax;
        };
        int k = 0;
        const int maxOps = e.fuseDepth > 0 ? std::min(64, e.fuseDepth) : 64;
        const int maxOrb = e.fuseDepth > 0 ? std::min(kOrbMaxOps, e.fuseDepth) : kOrbMaxOps;
        const bool useOrbits = e.sweepKind != kSweepTile;
        e.orbLds = e.sweepKind == kSweepOrbitLds;
        if (e.orbLds && (size_t)4 * e.nb * sizeof(double) > 65536) {
            HCHK(hipFuncSetAttribute((const void *)k_orbl<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)(4 * e.nb * sizeof(double))));
            HCHK(hipFuncSetAttribute((const void *)k_orbl<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)(4 * e.nb * sizeof(double))));
        }
        auto bitsA = [&](int kk) { return ((e.hops[kk].cre | e.hops[kk].ann) >> n) & amask; };
        auto bitsB = [&](int kk) { return (e.hops[kk].cre | e.hops[kk].ann) & amask; };
        while (k < M) {
            if (isDiag(k)) {
                k = diagRun(k);
                continue;
            }
            const uint32_t Uk = bitsA(k);
            if (useOrbits && __builtin_popcount(Uk) <= 2 && __builtin_popcount(bitsB(k)) <= 2) {
                // orbit segment: extend while both spins stay within two orbitals
                uint32_t Ua = Uk, Ub = bitsB(k);
                int k1 = k + 1;
                while (k1 < M && k1 - k < maxOrb && !isDiag(k1)) {
                    const uint32_t na2 = Ua | bitsA(k1), nb2 = Ub | bitsB(k1);
                    if (__builtin_popcount(na2) > 2 || __builtin_popcount(nb2) > 2) break;
                    Ua = na2;
                    Ub = nb2;
                    k1++;
                }
                Seg sg{k, k1, -1};
                sg.ua = Ua;
                sg.ub = Ub;
                sg.orbA = e.tileset_for(Ua, err, false);
                sg.orbB = e.tileset_for(Ub, err, true);
                if (sg.orbA < 0 || sg.orbB < 0) return false;
                sg.orbits = (long long)e.tilesets[sg.orbA]->ntiles * e.tilesets[sg.orbB]->ntiles;
                maxTiles = std::max(maxTiles, std::max(e.orb_grid(sg), e.tilesets[sg.orbA]->ntiles));
                e.segs.push_back(sg);
                k = k1;
                continue;
            }
            if (!fits(Uk)) {
                e.segs.push_back({k, k + 1, -1});
                e.ngeneric++;
                k++;
                continue;
            }
            uint32_t U = Uk;
            int k1 = k + 1;
            while (k1 < M && k1 - k < maxOps && !isDiag(k1)) {
                const uint32_t Un = U | (((e.hops[k1].cre | e.hops[k1].ann) >> n) & amask);
                if (!fits(Un)) break;
                U = Un;
                k1++;
            }
            const int ts = e.tileset_for(U, err);
            if (ts < 0) return false;
            maxTiles = std::max(maxTiles, e.tilesets[ts]->ntiles);
            e.segs.push_back({k, k1, ts});
            k = k1;
        }
        for (auto &t : e.tilesets) {
            const size_t lds = (size_t)2 * t->maxRows * e.nb * sizeof(double);
            if (lds > 65536) {
                HCHK(hipFuncSetAttribute((const void *)k_seg<2, true>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
                HCHK(hipFuncSetAttribute((const void *)k_seg<2, false>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
                HCHK(hipFuncSetAttribute((const void *)k_seg<1, false>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
            }
        }
    } else {
        for (int k = 0; k < M;) {
            if (isDiag(k)) {
                k = diagRun(k);
                continue;
            }
            e.segs.push_back({k, k + 1, -1});
            e.ngeneric++;
            k++;
        }
    }
    e.slots = std::max(maxTiles, e.Ggen);
    if (!e.build_orbit_programs(err)) return false;
    if (e.cplx) {
        // the lane layout (batches) and the two-group vector-major segments
        // are not used with complex amplitudes
        e.laneA.segs.clear();
        e.laneF.segs.clear();
        e.laneS.segs.clear();
        e.sweepS = e.laneSAllTwo = false;
        e.winOn = false;
        e.wpass.clear();
    } else if (!e.build_lane_segments(err)) {
        return false;
    }
    const int pl = e.planes();
    if (!e.d_cs.ensure((size_t)pl * M, err) || !e.d_csneg.ensure((size_t)pl * M, err) ||
        !e.d_partial.ensure((size_t)pl * M * e.slots, err) || !e.d_epart.ensure((size_t)pl * e.eslots_alloc(), err) ||
        !e.d_out.ensure((size_t)pl * (M + 1), err))
        return false;
    return true;
}

bool Engine::set_start(const std::vector<double> &start, std::string &err)
{
    EngineImpl &e = *d_;
    if ((long long)start.size() != e.N) {
        err = "start vector has the wrong size";
        return false;
    }
    e.drop_graphs();
    HCHK(hipMemcpy(e.d_start.p, start.data(), 8 * e.N, hipMemcpyHostToDevice));
    if (e.cplx) HCHK(hipMemset(e.d_start.p + e.N, 0, 8 * e.N));
    e.haveStart = true;
    e.hstart = start;
    e.startSym = e.symmetric_start();
    e.supDirty = e.hessDirty = true;
    return true;
}

bool Engine::set_complex(bool on, std::string &err)
{
    EngineImpl &e = *d_;
    if (e.N == 0) {
        err = "engine not initialised";
        return false;
    }
    if (on == e.cplx) return true;
    e.drop_graphs();
    e.cplx = on;
    e.haveStart = false;
    const size_t n = (size_t)e.planes() * e.N;
    if (!e.d_psi.ensure(n, err) || !e.d_phi.ensure(n, err) || !e.d_lam.ensure(n, err) || !e.d_start.ensure(n, err))
        return false;
    e.bnvec = 0;
    // re-segment: number operators become legal (or illegal) operators
    if (e.M > 0) return set_operators(e.ops4, std::vector<int>(e.order), std::vector<double>(e.scale), e.P, err);
    return true;
}

bool Engine::is_complex() const { return d_->cplx; }

bool Engine::set_start_complex(const std::vector<double> &re, const std::vector<double> &im, std::string &err)
{
    EngineImpl &e = *d_;
    if (!e.cplx) {
        err = "complex start state without complex amplitudes";
        return false;
    }
    if ((long long)re.size() != e.N || (long long)im.size() != e.N) {
        err = "start vector has the wrong size";
        return false;
    }
    e.drop_graphs();
    HCHK(hipMemcpy(e.d_start.p, re.data(), 8 * e.N, hipMemcpyHostToDevice));
    HCHK(hipMemcpy(e.d_start.p + e.N, im.data(), 8 * e.N, hipMemcpyHostToDevice));
    e.haveStart = true;
    e.hstart = re;
    e.startSym = false;  // complex amplitudes: no exchange-symmetric sigma
    e.supDirty = e.hessDirty = true;
    return true;
}

bool Engine::download_state_complex(double *re, double *im, std::string &err)
{
    EngineImpl &e = *d_;
    HCHK(hipMemcpyAsync(re, e.d_psi.p, 8 * e.N, hipMemcpyDeviceToHost, e.stream));
    if (e.cplx) {
        HCHK(hipMemcpyAsync(im, e.d_psi.p + e.N, 8 * e.N, hipMemcpyDeviceToHost, e.stream));
    } else {
        std::fill(im, im + e.N, 0.0);
    }
    HCHK(hipStreamSynchronize(e.stream));
    return true;
}

bool Engine::set_ham_rows(const std::vector<int64_t> &rowptr, const std::vector<int32_t> &col,
                          const std::vector<double> &val, std::string &err)
{
    EngineImpl &e = *d_;
    if ((long long)rowptr.size() != e.N + 1) {
        err = "row pointer has the wrong size";
        return false;
    }
    e.drop_graphs();
    if (!e.d_rowptr.ensure(rowptr.size(), err) || !e.d_col.ensure(col.size(), err) ||
        !e.d_val.ensure(val.size(), err))
        return false;
    HCHK(hipMemcpy(e.d_rowptr.p, rowptr.data(), 8 * rowptr.size(), hipMemcpyHostToDevice));
    if (!col.empty()) {
        HCHK(hipMemcpy(e.d_col.p, col.data(), 4 * col.size(), hipMemcpyHostToDevice));
        HCHK(hipMemcpy(e.d_val.p, val.data(), 8 * val.size(), hipMemcpyHostToDevice));
    }
    e.nnz = (long long)val.size();
    e.hkind = 1;
    return true;
}

bool Engine::set_ham_integrals(int norb, const std::vector<double> &h, const std::vector<double> &g, int storage,
                               std::string &err)
{
    EngineImpl &e = *d_;
    if (e.mode != kSz) {
        err = "integral Hamiltonians need the S_z sector (initial state and operators conserving S_z)";
        return false;
    }
    if (2 * norb != e.nq) {
        err = "number of qubits must be twice the number of orbitals";
        return false;
    }
    e.drop_graphs();
    const int n = e.n, ku = e.nup, kd = e.ndown;
    const long long W = 1 + (long long)kd * (n - kd) + (long long)ku * (n - ku) +
                        binom_ll(kd, 2) * binom_ll(n - kd, 2) + binom_ll(ku, 2) * binom_ll(n - ku, 2) +
                        (long long)ku * (n - ku) * kd * (n - kd);
    const size_t bytes = (size_t)W * e.N * 12;
    size_t freeB = 0, totalB = 0;
    HCHK(hipMemGetInfo(&freeB, &totalB));
    const bool fitsExplicit = bytes + (size_t)(4ull << 30) < freeB;
    if (!e.d_h.ensure(h.size(), err) || !e.d_g.ensure(g.size(), err)) return false;
    HCHK(hipMemcpy(e.d_h.p, h.data(), 8 * h.size(), hipMemcpyHostToDevice));
    HCHK(hipMemcpy(e.d_g.p, g.data(), 8 * g.size(), hipMemcpyHostToDevice));
    // automatic choice: matrix-free from 12 orbitals (its sigma outruns the
    // explicit rows there, single vectors and batches alike) or when the
    // rows do not fit; explicit rows below
    const bool preferMf = n >= 12;
    if (storage == kHamMatrixFree || (storage == kHamAuto && (!fitsExplicit || preferMf))) {
        // matrix-free: diagonal and same-spin single tables, then sigma on the fly
        const long long nsA = (long long)ku * (n - ku), nsB = (long long)kd * (n - kd);
        if (!e.d_diag.ensure(e.N, err) || !e.d_FA.ensure(std::max(1ll, nsA * e.na), err) ||
            !e.d_FB.ensure(std::max(1ll, nsB * e.nb), err))
            return false;
        if (e.nb > 16 * kBlock || n > 15) {
            err = "matrix-free sigma supports at most 15 orbitals";
            return false;
        }
        if (e.mf_lds() > 65536)
            HCHK(hipFuncSetAttribute((const void *)k_mf_sigma, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)e.mf_lds()));
        const MfArgs a = e.mfargs();
        const long long cnt = std::max<long long>(e.N, std::max(e.na, e.nb));
        hipLaunchKernelGGL(k_mf_tables, dim3((unsigned)((cnt + kBlock - 1) / kBlock)), dim3(kBlock), 0, e.stream, a);
        LCHK();
        HCHK(hipStreamSynchronize(e.stream));
        e.d_col.release();
        e.d_val.release();
        e.d_rowptr.release();
        e.W = 0;
        e.nnz = W * e.N;  // implicit entries
        e.hkind = 3;
        e.sigv = false;
        e.sig1 = false;
        if (n <= 14 && e.nb <= kSig1MaxNb) {
            // tables of the split sigma: dense same-spin rows, beta entries
            const size_t na = e.na, nb = e.nb;
            e.ntp = (kd * (n - kd) + kd + 3) / 4 * 4;
            e.ntp16 = (kd * (n - kd) + 3) / 4 * 4;
            if (!e.d_Vaa.ensure(na * na, err) || !e.d_Vbb.ensure(nb * nb, err) ||
                !e.d_ts.ensure(nb * kSigWNtp, err) || !e.d_ts1.ensure(nb * e.ntp, err))
                return false;
            HCHK(hipMemsetAsync(e.d_Vaa.p, 0, na * na * sizeof(double), e.stream));
            HCHK(hipMemsetAsync(e.d_Vbb.p, 0, nb * nb * sizeof(double), e.stream));
            hipLaunchKernelGGL(k_dense_spin, dim3((unsigned)((na + kBlock - 1) / kBlock)), dim3(kBlock), 0, e.stream,
                               a, 1, e.d_Vaa.p);
            LCHK();
            hipLaunchKernelGGL(k_dense_spin, dim3((unsigned)((nb + kBlock - 1) / kBlock)), dim3(kBlock), 0, e.stream,
                               a, 0, e.d_Vbb.p);
            LCHK();
            const bool lanesOk = n <= 12 && e.nb <= kSigMaxNb && e.ntp16 >= 4 && e.ntp16 <= 36;
            if (lanesOk) {
                hipLaunchKernelGGL(k_beta_table, dim3((unsigned)((nb + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                                   e.stream, a, e.d_ts.p, kSigWNtp, kSigWVC * 8, kSigNN, 17, 1);
                LCHK();
            }
            hipLaunchKernelGGL(k_beta_table, dim3((unsigned)((nb + kBlock - 1) / kBlock)), dim3(kBlock), 0, e.stream,
                               a, e.d_ts1.p, e.ntp, 8, kSig1NN, 16, 0);
            LCHK();
            HCHK(hipStreamSynchronize(e.stream));
            e.vSame = false;
            if (na == nb) {
                std::vector<double> va(na * na), vb(nb * nb);
                HCHK(hipMemcpy(va.data(), e.d_Vaa.p, va.size() * sizeof(double), hipMemcpyDeviceToHost));
                HCHK(hipMemcpy(vb.data(), e.d_Vbb.p, vb.size() * sizeof(double), hipMemcpyDeviceToHost));
                e.vSame = va == vb;
            }
            if (lanesOk)
                HCHK(hipFuncSetAttribute((const void *)sigw_kernel(e.ntp16 / 4),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)kSigWLds));
            e.sigh = lanesOk && n == 2 * kd && sigh_kernel((kd - 1) * (n - kd), n - kd) != nullptr;
            e.sigg = lanesOk && n == 12 && kd == 6 && e.nb == 924;
            if (e.sigg) {
                HCHK(hipFuncSetAttribute((const void *)k_sigg_sparse<6>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                         (int)kSigWLds));
                // source rows per alpha row in ascending rank: concurrently
                // running workgroups with nearby targets stage nearby (often
                // the same) source rows at about the same time
                const int ku = __builtin_popcount(e.unA[0]);
                const int R = 1 + ku * (n - ku);
                std::vector<int4> src((size_t)na * R);
                for (size_t ar = 0; ar < na; ar++) {
                    const uint32_t al = e.unA[ar], full = (1u << n) - 1;
                    std::vector<int4> row;
                    row.push_back(make_int4((int)ar, -1, -1, 0));
                    for (uint32_t o = al; o; o &= o - 1)
                        for (uint32_t w = ~al & full; w; w &= w - 1) {
                            const int rA = __builtin_ctz(o), qA = __builtin_ctz(w);
                            row.push_back(make_int4(e.rkA[al ^ (1u << rA) ^ (1u << qA)], rA, qA, 0));
                        }
                    std::sort(row.begin(), row.end(), [](const int4 &p, const int4 &q) { return p.x < q.x; });
                    std::copy(row.begin(), row.end(), src.begin() + ar * R);
                }
                if (!e.d_sigsrc.ensure(src.size(), err)) return false;
                HCHK(hipMemcpy(e.d_sigsrc.p, src.data(), src.size() * sizeof(int4), hipMemcpyHostToDevice));
            }
            if (e.sigh && !e.build_sigh_tables(err)) return false;
            HCHK(hipFuncSetAttribute((const void *)k_sig1_sparse, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)kSig1Lds));
            {
                const void *fns[4] = {(const void *)k_sig1r<2>, (const void *)k_sig1r<4>,
                                      (const void *)k_sig1r<6>, (const void *)k_sig1r<8>};
                for (int q = 0; q < 4; q++) {
                    const size_t lds = (size_t)(2 * q + 2) * (kSig1GS + e.nb) * sizeof(double);
                    if (lds <= e.ldsDevice)
                        HCHK(hipFuncSetAttribute(fns[q], hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
                }
            }
            if (!e.blas) {
                if (rocblas_create_handle(&e.blas) != rocblas_status_success) {
                    e.blas = nullptr;
                    err = "rocblas_create_handle failed";
                    return false;
                }
            }
            if (rocblas_set_stream(e.blas, e.stream) != rocblas_status_success) {
                err = "rocblas_set_stream failed";
                return false;
            }
            e.sigv = lanesOk;
            e.sig1 = true;
        }
        return true;
    }
    if (!fitsExplicit) {
        err = "explicit Hamiltonian rows do not fit in device memory";
        return false;
    }
    const DBuf<double> &dh = e.d_h, &dg = e.d_g;
    e.d_rowptr.release();
    if (!e.d_col.ensure((size_t)W * e.N, err) || !e.d_val.ensure((size_t)W * e.N, err)) return false;
    EllBuildArgs a;
    a.n = n;
    a.ku = ku;
    a.kd = kd;
    a.nb = e.nb;
    a.N = e.N;
    a.W = (int)W;
    a.unA = e.d_unA.p;
    a.unB = e.d_unB.p;
    a.rkA = e.d_rkA.p;
    a.rkB = e.d_rkB.p;
    a.h = dh.p;
    a.g = dg.p;
    a.col = e.d_col.p;
    a.val = e.d_val.p;
    hipLaunchKernelGGL(k_ell_build, dim3((unsigned)((e.N + kBlock - 1) / kBlock)), dim3(kBlock), 0, e.stream, a);
    LCHK();
    HCHK(hipStreamSynchronize(e.stream));
    e.W = (int)W;
    e.nnz = W * e.N;
    e.hkind = 2;
    return true;
}

bool Engine::evolve(const double *angles, std::string &err)
{
    EngineImpl &e = *d_;
    if (!e.haveStart || e.M == 0) {
        err = "engine not set up";
        return false;
    }
    const int pl = e.planes();
    if (!e.upload_cs_single(angles, err)) return false;
    HCHK(hipMemcpyAsync(e.d_psi.p, e.d_start.p, 8 * e.N * pl, hipMemcpyDeviceToDevice, e.stream));
    if (!e.sweep_forward(e.d_psi.p, pl, e.d_cs.p, err)) return false;
    HCHK(hipStreamSynchronize(e.stream));
    return true;
}

bool Engine::energy(double *E, std::string &err)
{
    EngineImpl &e = *d_;
    const int pl = e.planes(), es = e.eslots();
    if (!e.ham_apply(e.d_psi.p, e.d_lam.p, pl, e.d_epart.p, err)) return false;
    hipLaunchKernelGGL(k_reduce, dim3(pl), dim3(kBlock), 0, e.stream, e.d_epart.p, pl, es, (long long)es, 1LL,
                       e.d_out.p, 1.0);
    LCHK();
    double parts[2] = {0, 0};
    HCHK(hipMemcpyAsync(parts, e.d_out.p, 8 * pl, hipMemcpyDeviceToHost, e.stream));
    HCHK(hipStreamSynchronize(e.stream));
    *E = parts[0] + parts[1];  // real plane, imaginary plane
    return true;
}

bool Engine::energy_gradient(const double *angles, double *E, double *deriv, std::string &err)
{
    EngineImpl &e = *d_;
    if (!e.haveStart || e.M == 0 || e.hkind == 0) {
        err = "engine not set up";
        return false;
    }
    if (!e.upload_cs_single(angles, err)) return false;
    const bool sym = e.vec_symmetric(angles);
    // complex amplitudes: both planes through every step; out holds the
    // per-operator brackets of each plane, then each plane's energy
    const int pl = e.planes(), es = e.eslots();
    const long long NP = e.N * pl;
    const unsigned cgrid = (unsigned)std::min<long long>((NP + kBlock - 1) / kBlock, 4096);
    auto enqueue = [&]() -> bool {
        hipLaunchKernelGGL(k_bcast, dim3(cgrid), dim3(kBlock), 0, e.stream, e.d_start.p, e.d_psi.p, NP, 1);
        LCHK();
        if (!e.sweep_forward(e.d_psi.p, pl, e.d_cs.p, err)) return false;
        hipLaunchKernelGGL(k_bcast, dim3(cgrid), dim3(kBlock), 0, e.stream, e.d_psi.p, e.d_phi.p, NP, 1);
        LCHK();
        if (!e.ham_apply(e.d_phi.p, e.d_lam.p, pl, e.d_epart.p, err, sym)) return false;
        hipLaunchKernelGGL(k_reduce, dim3(pl), dim3(kBlock), 0, e.stream, e.d_epart.p, pl, es, (long long)es, 1LL,
                           e.d_out.p + (size_t)pl * e.M, 1.0);
        LCHK();
        if (!e.sweep_reverse(e.d_phi.p, e.d_lam.p, pl, e.d_csneg.p, e.d_partial.p, err)) return false;
        hipLaunchKernelGGL(k_reduce, dim3(pl * e.M), dim3(kBlock), 0, e.stream, e.d_partial.p, pl * e.M, e.slots,
                           (long long)e.slots, 1LL, e.d_out.p, 2.0);
        LCHK();
        return true;
    };
    if (e.useGraph && !e.timingOn) {
        if (e.egGraph && (e.egGraphPtrs != e.eg_ptrs() || e.egSym != sym)) {
            (void)hipGraphExecDestroy(e.egGraph);
            e.egGraph = nullptr;
        }
        if (!e.egGraph) {
            // captured once: the sin/cos tables and every buffer keep their
            // device addresses, so the graph replays for any angles.  One
            // eager run first: library GEMM kernels load lazily and must not
            // load under capture.
            if (!enqueue()) return false;
            HCHK(hipStreamSynchronize(e.stream));
            hipGraph_t g = nullptr;
            HCHK(hipStreamBeginCapture(e.stream, hipStreamCaptureModeThreadLocal));
            const bool ok = enqueue();
            const hipError_t ec = hipStreamEndCapture(e.stream, &g);
            if (!ok) {
                if (g) (void)hipGraphDestroy(g);
                return false;
            }
            HCHK(ec);
            const hipError_t ei = hipGraphInstantiate(&e.egGraph, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            HCHK(ei);
            e.egGraphPtrs = e.eg_ptrs();
            e.egSym = sym;
        }
        HCHK(hipGraphLaunch(e.egGraph, e.stream));
    } else if (!enqueue()) {
        return false;
    }
    return e.collect_single(deriv, E, err);
}

bool Engine::gradient(const double *angles, double *E, double *deriv, std::string &err)
{
    EngineImpl &e = *d_;
    // sin/cos of the current angles (the forward sweep may have used them
    // already; recomputing keeps this call self-contained)
    if (!e.upload_cs_single(angles, err)) return false;
    const int pl = e.planes(), es = e.eslots();
    HCHK(hipMemcpyAsync(e.d_phi.p, e.d_psi.p, 8 * e.N * pl, hipMemcpyDeviceToDevice, e.stream));
    if (!e.ham_apply(e.d_phi.p, e.d_lam.p, pl, e.d_epart.p, err)) return false;
    hipLaunchKernelGGL(k_reduce, dim3(pl), dim3(kBlock), 0, e.stream, e.d_epart.p, pl, es, (long long)es, 1LL,
                       e.d_out.p + (size_t)pl * e.M, 1.0);
    LCHK();
    if (!e.sweep_reverse(e.d_phi.p, e.d_lam.p, pl, e.d_csneg.p, e.d_partial.p, err)) return false;
    hipLaunchKernelGGL(k_reduce, dim3(pl * e.M), dim3(kBlock), 0, e.stream, e.d_partial.p, pl * e.M, e.slots,
                       (long long)e.slots, 1LL, e.d_out.p, 2.0);
    LCHK();
    return e.collect_single(deriv, E, err);
}

bool Engine::overlap_gradient(const double *angles, const double *proj, double *overlap, double *deriv,
                              std::string &err)
{
    EngineImpl &e = *d_;
    if (!e.real_only("the overlap gradient", err)) return false;
    if (!e.haveStart || e.M == 0) {
        err = "engine not set up";
        return false;
    }
    if (!e.upload_cs(angles, 1, e.d_cs, e.d_csneg, err)) return false;
    HCHK(hipMemcpyAsync(e.d_psi.p, e.d_start.p, 8 * e.N, hipMemcpyDeviceToDevice, e.stream));
    if (!e.sweep_forward(e.d_psi.p, 1, e.d_cs.p, err)) return false;
    HCHK(hipMemcpyAsync(e.d_phi.p, e.d_psi.p, 8 * e.N, hipMemcpyDeviceToDevice, e.stream));
    HCHK(hipMemcpyAsync(e.d_lam.p, proj, 8 * e.N, hipMemcpyHostToDevice, e.stream));
    // overlap partials in the energy partial rows (eslots_alloc() >= this grid)
    const int G = (int)((e.N + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_dot_rows, dim3(G, 1), dim3(kBlock), 0, e.stream, e.d_phi.p, e.d_lam.p, e.N, e.N,
                       e.d_epart.p);
    LCHK();
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(kBlock), 0, e.stream, e.d_epart.p, 1, G, 0LL, 1LL, e.d_out.p + e.M,
                       1.0);
    LCHK();
    if (!e.sweep_reverse(e.d_phi.p, e.d_lam.p, 1, e.d_csneg.p, e.d_partial.p, err)) return false;
    // brackets without the factor 2 of the energy gradient (src/fusedevolve.cpp:2882-2884)
    hipLaunchKernelGGL(k_reduce, dim3(e.M), dim3(kBlock), 0, e.stream, e.d_partial.p, e.M, e.slots,
                       (long long)e.slots, 1LL, e.d_out.p, 1.0);
    LCHK();
    std::vector<double> out(e.M + 1);
    HCHK(hipMemcpyAsync(out.data(), e.d_out.p, 8 * (e.M + 1), hipMemcpyDeviceToHost, e.stream));
    HCHK(hipStreamSynchronize(e.stream));
    std::memcpy(deriv, out.data(), 8 * e.M);
    if (overlap) *overlap = out[e.M];
    return true;
}

bool Engine::download_state(double *out, std::string &err)
{
    EngineImpl &e = *d_;
    HCHK(hipMemcpyAsync(out, e.d_psi.p, 8 * e.N, hipMemcpyDeviceToHost, e.stream));
    HCHK(hipStreamSynchronize(e.stream));
    return true;
}

bool Engine::download_ham_state(double *out, std::string &err)
{
    EngineImpl &e = *d_;
    if (!e.real_only("H psi", err)) return false;
    if (!e.ham_apply(e.d_psi.p, e.d_lam.p, 1, e.d_epart.p, err)) return false;
    HCHK(hipMemcpyAsync(out, e.d_lam.p, 8 * e.N, hipMemcpyDeviceToHost, e.stream));
    HCHK(hipStreamSynchronize(e.stream));
    return true;
}

// The device address of `scratch` on the current device, looked up once per
// device (the symbol's host-side name is no device address).
static bool noalloc_scratch(double **out, std::string &err)
{
    static std::mutex mu;
    static std::map<int, double *> ptrs;
    int dev = 0;
    HCHK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    auto it = ptrs.find(dev);
    if (it == ptrs.end()) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            err = std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e);
            return false;
        }
        it = ptrs.emplace(dev, static_cast<double *>(p)).first;
    }
    *out = it->second;
    return true;
}

size_t Engine::noalloc_scratch_doubles() { return kNoAllocScratchDoubles; }

size_t Engine::noalloc_scratch_need() const
{
    const EngineImpl &e = *d_;
    const size_t M = e.M;
    auto r = [](size_t n) { return (n + 31) & ~(size_t)31; };  // 256-byte aligned parts
    return r(2 * M) + r(2 * M) + r(M + 1) + r(e.eslots_alloc()) + r(M * (size_t)e.slots);
}

// energy_gradient without graph, rocBLAS or device allocation: the angle
// tables, energy partials, bracket partials and results live in `scratch`
// (layout: cos/sin 2M | cos/-sin 2M | results M+1 | energy partials |
// bracket partials M x slots), the states in the engine's own buffers.
// The sigma is the explicit rows or the all-in-one matrix-free kernel.
bool Engine::energy_gradient_noalloc(const double *angles, double *E, double *deriv, std::string &err)
{
    EngineImpl &e = *d_;
    if (!e.real_only("the no-allocation evaluation", err)) return false;
    if (!e.haveStart || e.M == 0 || e.hkind == 0) {
        err = "engine not set up";
        return false;
    }
    const size_t need = noalloc_scratch_need();
    if (need > kNoAllocScratchDoubles) {
        err = "this set-up needs " + std::to_string(need) + " doubles of scratch; the no-allocation entry point has " +
              std::to_string(kNoAllocScratchDoubles);
        return false;
    }
    double *base = nullptr;
    if (!noalloc_scratch(&base, err)) return false;
    const size_t M = e.M;
    auto r = [](size_t n) { return (n + 31) & ~(size_t)31; };
    double2 *cs = reinterpret_cast<double2 *>(base);
    double2 *csneg = reinterpret_cast<double2 *>(base + r(2 * M));
    double *out = base + 2 * r(2 * M);
    double *epart = out + r(M + 1);
    double *partial = epart + r(e.eslots_alloc());
    e.fill_cs(angles, 1);
    e.h_csneg.resize(M);
    for (size_t k = 0; k < M; k++) e.h_csneg[k] = make_double2(e.h_cs[k].x, -e.h_cs[k].y);
    HCHK(hipMemcpyAsync(cs, e.h_cs.data(), M * sizeof(double2), hipMemcpyHostToDevice, e.stream));
    HCHK(hipMemcpyAsync(csneg, e.h_csneg.data(), M * sizeof(double2), hipMemcpyHostToDevice, e.stream));
    const unsigned cgrid = (unsigned)std::min<long long>((e.N + kBlock - 1) / kBlock, 4096);
    hipLaunchKernelGGL(k_bcast, dim3(cgrid), dim3(kBlock), 0, e.stream, e.d_start.p, e.d_psi.p, e.N, 1);
    LCHK();
    if (!e.sweep_forward(e.d_psi.p, 1, cs, err)) return false;
    hipLaunchKernelGGL(k_bcast, dim3(cgrid), dim3(kBlock), 0, e.stream, e.d_psi.p, e.d_phi.p, e.N, 1);
    LCHK();
    if (e.hkind == 3) {
        hipLaunchKernelGGL(k_mf_sigma, dim3(e.na, 1), dim3(kBlock), e.mf_lds(), e.stream, e.mfargs(), e.d_phi.p,
                           e.d_lam.p, e.N, epart);
        LCHK();
    } else if (!e.ham_apply(e.d_phi.p, e.d_lam.p, 1, epart, err)) {
        return false;
    }
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(kBlock), 0, e.stream, epart, 1, e.eslots(), 0LL, 1LL, out + M, 1.0);
    LCHK();
    // sweep_reverse clears the bracket partials before it accumulates
    if (!e.sweep_reverse(e.d_phi.p, e.d_lam.p, 1, csneg, partial, err)) return false;
    hipLaunchKernelGGL(k_reduce, dim3((unsigned)M), dim3(kBlock), 0, e.stream, partial, (int)M, e.slots,
                       (long long)e.slots, 1LL, out, 2.0);
    LCHK();
    std::vector<double> res(M + 1);
    HCHK(hipMemcpyAsync(res.data(), out, 8 * (M + 1), hipMemcpyDeviceToHost, e.stream));
    HCHK(hipStreamSynchronize(e.stream));
    std::memcpy(deriv, res.data(), 8 * M);
    if (E) *E = res[M];
    return true;
}

bool Engine::apply_ham(const double *x, double *y, std::string &err)
{
    EngineImpl &e = *d_;
    if (e.hkind == 0) {
        err = "no Hamiltonian";
        return false;
    }
    HCHK(hipMemcpyAsync(e.d_phi.p, x, 8 * e.N, hipMemcpyHostToDevice, e.stream));
    if (!e.ham_apply(e.d_phi.p, e.d_lam.p, 1, e.d_epart.p, err)) return false;
    HCHK(hipMemcpyAsync(y, e.d_lam.p, 8 * e.N, hipMemcpyDeviceToHost, e.stream));
    HCHK(hipStreamSynchronize(e.stream));
    return true;
}

bool Engine::subspace(int nvec, const double *angles, double *Hmat, double *Smat, std::string &err)
{
    EngineImpl &e = *d_;
    if (!e.real_only("subspace diagonalisation", err)) return false;
    if (!e.haveStart || e.M == 0 || e.hkind == 0) {
        err = "engine not set up";
        return false;
    }
    if (nvec < 1) {
        err = "empty set of angle vectors";
        return false;
    }
    const long long N = e.N;
    const int chunks = 64;
    const size_t nn = (size_t)nvec * nvec;
    if (!e.hz_B.ensure((size_t)nvec * N, err) || !e.hz_HB.ensure((size_t)nvec * N, err) ||
        !e.hz_epart.ensure((size_t)nvec * e.eslots_alloc(), err) || !e.hz_gram.ensure(chunks * nn, err) ||
        !e.hz_out.ensure(2 * nn, err))
        return false;
    if (!e.upload_cs(angles, nvec, e.sub_cs, e.sub_csneg, err)) return false;
    const long long total = (long long)nvec * N;
    hipLaunchKernelGGL(k_bcast, dim3((unsigned)std::min<long long>((total + kBlock - 1) / kBlock, 16384)),
                       dim3(kBlock), 0, e.stream, e.d_start.p, e.hz_B.p, N, nvec);
    LCHK();
    if (!e.sweep_forward(e.hz_B.p, nvec, e.sub_cs.p, err)) return false;
    if (!e.ham_apply(e.hz_B.p, e.hz_HB.p, nvec, e.hz_epart.p, err)) return false;
    const int tiles = (nvec + 15) / 16;
    for (int which = 0; which < 2; which++) {
        hipLaunchKernelGGL(k_gram, dim3(tiles, tiles, chunks), dim3(kBlock), 0, e.stream, e.hz_B.p,
                           which ? e.hz_B.p : e.hz_HB.p, N, nvec, N, chunks, e.hz_gram.p);
        LCHK();
        hipLaunchKernelGGL(k_reduce, dim3((unsigned)nn), dim3(kBlock), 0, e.stream, e.hz_gram.p, (int)nn, chunks,
                           1LL, (long long)nn, e.hz_out.p + which * nn, 1.0);
        LCHK();
    }
    std::vector<double> out(2 * nn);
    HCHK(hipMemcpyAsync(out.data(), e.hz_out.p, 8 * out.size(), hipMemcpyDeviceToHost, e.stream));
    HCHK(hipStreamSynchronize(e.stream));
    std::memcpy(Hmat, out.data(), 8 * nn);
    std::memcpy(Smat, out.data() + nn, 8 * nn);
    return true;
}

bool Engine::batch_upload(int nvec, const double *angles, std::string &err)
{
    EngineImpl &e = *d_;
    if (!e.real_only("a batch of angle vectors", err)) return false;
    if (!e.haveStart || e.M == 0 || e.hkind == 0) {
        err = "engine not set up";
        return false;
    }
    if (nvec < 1) {
        err = "empty batch";
        return false;
    }
    const size_t NV = (size_t)nvec * e.N;
    if (!e.b_psi.ensure(NV, err) || !e.b_lam.ensure(NV, err) ||
        !e.b_epart.ensure((size_t)nvec * e.eslots_alloc(), err) || !e.b_out.ensure((size_t)nvec * (e.M + 1), err))
        return false;
    if (e.use_lanes(nvec)) {
        const size_t nch = (size_t)(nvec + kLanes - 1) / kLanes;
        if (!e.bl_psi.ensure(nch * e.N * kLanes, err) || !e.bl_lam.ensure(nch * e.N * kLanes, err) ||
            !e.bl_partial.ensure(nch * e.M * e.lslots * kLanes, err) || !e.bl_epart.ensure(nch * 512 * kLanes, err) ||
            !e.bl_wep.ensure((size_t)nch * (std::max(1, e.win_erows()) + 16) * kLanes, err))
            return false;
        const bool sym = e.batch_symmetric(angles, nvec);
        if (sym && !e.bl_z.ensure((size_t)nch * e.N * kLanes, err)) return false;
        if (sym != e.bSym) e.drop_graphs();
        e.bSym = sym;
        // stored forward states 0 .. segments, when they fit in half the
        // free memory
        bool ck = e.ckOpt && (e.ckShape || e.winOn) && e.hkind == 3 && e.sigv;
        const double *ckWas = e.bl_ck.p;
        if (ck) {
            const size_t need = (size_t)(e.nfwd() + 1) * nch * (size_t)e.N * kLanes;
            size_t freeB = 0, totalB = 0;
            HCHK(hipMemGetInfo(&freeB, &totalB));
            const size_t have = e.bl_ck.n;
            ck = need <= have || (need - have) * sizeof(double) <= freeB / 2;
            if (ck && !e.bl_ck.ensure(need, err)) return false;
        }
        if (ck != e.ckOn || e.bl_ck.p != ckWas) e.drop_graphs();
        e.ckOn = ck;
        if (e.supDirty || e.supCk != ck) {
            if (!e.compute_support(ck, err)) return false;
            e.drop_graphs();
        }
        if (ck && (e.ckZeroP != e.bl_ck.p || e.ckZeroNch != nch || e.ckZeroGen != e.supGen)) {
            // the forward sweep writes only the reachable part of each state
            HCHK(hipMemsetAsync(e.bl_ck.p, 0, e.bl_ck.n * sizeof(double), e.stream));
            HCHK(hipStreamSynchronize(e.stream));
            e.ckZeroP = e.bl_ck.p;
            e.ckZeroNch = nch;
            e.ckZeroGen = e.supGen;
        }
        if (!e.upload_cs_lanes(angles, nvec, err)) return false;
    } else {
        e.bSym = false;
        bool symv = true;
        for (int v = 0; v < nvec && symv; v++) symv = e.vec_symmetric(angles + (size_t)v * e.M);
        if (symv != e.bSymV) e.drop_graphs();
        e.bSymV = symv;
        if (!e.b_partial.ensure((size_t)nvec * e.M * e.slots, err)) return false;
        if (!e.upload_cs(angles, nvec, e.b_cs, e.b_csneg, err)) return false;
    }
    std::vector<const void *> ptrs = e.batch_ptrs();
    if (ptrs != e.uploadPtrs) {
        e.drop_graphs();
        e.uploadPtrs.swap(ptrs);
    }
    e.bnvec = nvec;
    return true;
}

// Enqueues psi_v = U(x_v) psi_0, E_v, and (optionally) the adjoint sweep for
// every vector of the uploaded batch; results stay in b_out.  Batches of at
// least laneMin vectors sweep in lane layout and transpose around the
// Hamiltonian application.  The sequence is captured once per (batch size,
// gradient, layout) and replayed as one hipGraph launch.
// The batch sequence for the uploaded batch on the engine's stream (no graph).
bool Engine::batch_enqueue(bool with_gradient, std::string &err)
{
    EngineImpl &e = *d_;
    const int nvec = e.bnvec;
    if (nvec < 1) {
        err = "no angle batch uploaded";
        return false;
    }
    const bool lanes = e.use_lanes(nvec);
    const int es = e.eslots();
    if (lanes) {
        const int nch = (nvec + kLanes - 1) / kLanes;
        const long long total = (long long)nch * e.N * kLanes;
        const dim3 tgrid((unsigned)((e.N + kLanes - 1) / kLanes), nch);
        double *psi0 = e.ckOn ? e.ck_state(0, nch) : e.bl_psi.p;
        if (!e.first_reads_start()) {
            hipLaunchKernelGGL(k_bcast_lanes,
                               dim3((unsigned)std::min<long long>((total + kBlock - 1) / kBlock, 16384)), dim3(kBlock),
                               0, e.stream, e.d_start.p, psi0, e.N, nch);
            LCHK();
        }
        if (!e.sweep_forward_lanes(psi0, nch, err)) return false;
        if (e.hkind == 3 && e.sigv) {
            // matrix-free sigma in lane layout: no transposes
            const double *psi = e.ckOn ? e.ck_state(e.nfwd(), nch) : e.bl_psi.p;
            // with the gradient, window passes in registers take the energies
            // in their first adjoint pass (its tiles hold the final state and
            // H psi) instead of a separate dot over both vectors
            const bool fe = with_gradient && e.fused_energy() && e.winFuseE;
            if (!e.ham_apply_lanes(psi, e.bl_lam.p, nch, nvec, e.b_out.p, err, nullptr, !fe,
                                   e.bSym ? e.bl_z.p : nullptr))
                return false;
            if (with_gradient) {
                if (!e.sweep_reverse_lanes(e.bl_psi.p, e.bl_lam.p, nch, err, fe ? e.bl_wep.p : nullptr)) return false;
                hipLaunchKernelGGL(k_reduce_lanes, dim3(nch * e.M), dim3(kBlock), 0, e.stream, e.bl_partial.p,
                                   e.M, e.lslots, e.d_lrows.p, nvec, e.b_out.p + nvec, 2.0);
                LCHK();
                if (fe) {
                    // per row group, then over the groups
                    const int P = e.win_egroups();
                    double *tmp = e.bl_wep.p + (size_t)nch * e.win_erows() * kLanes;
                    hipLaunchKernelGGL(k_reduce_lanes, dim3(nch * P), dim3(kBlock), 0, e.stream, e.bl_wep.p, P,
                                       e.win_erows() / P, nullptr, nvec, tmp, 1.0);
                    LCHK();
                    hipLaunchKernelGGL(k_reduce, dim3(nvec), dim3(kBlock), 0, e.stream, tmp, nvec, P, (long long)P,
                                       1LL, e.b_out.p, 1.0);
                    LCHK();
                }
            }
            return true;
        }
        hipLaunchKernelGGL(k_from_lanes, tgrid, dim3(kBlock), 0, e.stream, e.bl_psi.p, e.b_psi.p, e.N, nvec);
        LCHK();
        if (!e.ham_apply(e.b_psi.p, e.b_lam.p, nvec, e.b_epart.p, err)) return false;
        hipLaunchKernelGGL(k_reduce, dim3(nvec), dim3(kBlock), 0, e.stream, e.b_epart.p, nvec, es, (long long)es,
                           1LL, e.b_out.p, 1.0);
        LCHK();
        if (with_gradient) {
            hipLaunchKernelGGL(k_to_lanes, tgrid, dim3(kBlock), 0, e.stream, e.b_lam.p, e.bl_lam.p, e.N, nvec);
            LCHK();
            if (!e.sweep_reverse_lanes(e.bl_psi.p, e.bl_lam.p, nch, err)) return false;
            hipLaunchKernelGGL(k_reduce_lanes, dim3(nch * e.M), dim3(kBlock), 0, e.stream, e.bl_partial.p, e.M,
                               e.lslots, e.d_lrows.p, nvec, e.b_out.p + nvec, 2.0);
            LCHK();
        }
        return true;
    }
    const long long total = (long long)nvec * e.N;
    hipLaunchKernelGGL(k_bcast, dim3((unsigned)std::min<long long>((total + kBlock - 1) / kBlock, 16384)),
                       dim3(kBlock), 0, e.stream, e.d_start.p, e.b_psi.p, e.N, nvec);
    LCHK();
    if (!e.sweep_forward(e.b_psi.p, nvec, e.b_cs.p, err)) return false;
    if (!e.ham_apply(e.b_psi.p, e.b_lam.p, nvec, e.b_epart.p, err, e.bSymV)) return false;
    hipLaunchKernelGGL(k_reduce, dim3(nvec), dim3(kBlock), 0, e.stream, e.b_epart.p, nvec, es, (long long)es, 1LL,
                       e.b_out.p, 1.0);
    LCHK();
    if (with_gradient) {
        if (!e.sweep_reverse(e.b_psi.p, e.b_lam.p, nvec, e.b_csneg.p, e.b_partial.p, err)) return false;
        hipLaunchKernelGGL(k_reduce, dim3(nvec * e.M), dim3(kBlock), 0, e.stream, e.b_partial.p, nvec * e.M,
                           e.slots, (long long)e.slots, 1LL, e.b_out.p + nvec, 2.0);
        LCHK();
    }
    return true;
}

bool Engine::batch_run(bool with_gradient, std::string &err)
{
    EngineImpl &e = *d_;
    const int nvec = e.bnvec;
    if (nvec < 1) {
        err = "no angle batch uploaded";
        return false;
    }
    const bool lanes = e.use_lanes(nvec);
    auto enqueue = [&]() -> bool { return batch_enqueue(with_gradient, err); };
    if (!e.useGraph || e.timingOn) return enqueue();
    if (e.bGraph && (e.bGraphNvec != nvec || e.bGraphGrad != with_gradient || e.bGraphLanes != lanes ||
                     e.bGraphPtrs != e.batch_ptrs())) {
        (void)hipGraphExecDestroy(e.bGraph);
        e.bGraph = nullptr;
    }
    if (!e.bGraph) {
        // the graph is keyed on (size, gradient, layout) and on the buffer
        // addresses it captures.  One eager run first: library GEMM kernels
        // load lazily and must not load under capture.
        if (!enqueue()) return false;
        HCHK(hipStreamSynchronize(e.stream));
        hipGraph_t g = nullptr;
        HCHK(hipStreamBeginCapture(e.stream, hipStreamCaptureModeThreadLocal));
        const bool ok = enqueue();
        const hipError_t ec = hipStreamEndCapture(e.stream, &g);
        if (!ok) {
            if (g) (void)hipGraphDestroy(g);
            return false;
        }
        HCHK(ec);
        const hipError_t ei = hipGraphInstantiate(&e.bGraph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HCHK(ei);
        e.bGraphNvec = nvec;
        e.bGraphGrad = with_gradient;
        e.bGraphLanes = lanes;
        e.bGraphPtrs = e.batch_ptrs();
    }
    HCHK(hipGraphLaunch(e.bGraph, e.stream));
    return true;
}

bool Engine::batch_download(double *E, double *deriv, std::string &err)
{
    EngineImpl &e = *d_;
    const int nvec = e.bnvec;
    HCHK(hipMemcpyAsync(E, e.b_out.p, 8 * nvec, hipMemcpyDeviceToHost, e.stream));
    if (deriv)
        HCHK(hipMemcpyAsync(deriv, e.b_out.p + nvec, 8 * (size_t)nvec * e.M, hipMemcpyDeviceToHost, e.stream));
    HCHK(hipStreamSynchronize(e.stream));
    e.collect_marks();
    return true;
}

int Engine::batch_size() const { return d_->bnvec; }

void *Engine::stream() const { return (void *)d_->stream; }

bool Engine::timing_on() const { return d_->timingOn; }

int Engine::generation() const { return d_->generation; }

bool Engine::set_option(int key, int value, std::string &err)
{
    EngineImpl &e = *d_;
    switch (key) {
    case kOptSweep:
        if (value < kSweepOrbit || value > kSweepOrbit2) {
            err = "unknown sweep kind";
            return false;
        }
        if (value == e.sweepKind) return true;
        e.sweepKind = value;
        if (e.M > 0) {
            const std::vector<int8_t> ops4 = e.ops4;
            const std::vector<int> order = e.order;
            const std::vector<double> scale = e.scale;
            if (!set_operators(ops4, order, scale, e.P, err)) return false;
        }
        e.bnvec = 0;  // batch workspaces are re-sized for the new segmentation at the next upload
        return true;
    case kOptLaneMin:
        if (value < 0) {
            err = "lane threshold must be >= 0";
            return false;
        }
        e.laneMin = value;
        e.bnvec = 0;  // the uploaded batch's tables are in the old layout
        e.drop_graphs();
        return true;
    case kOptGraph:
        e.useGraph = value != 0;
        e.drop_graphs();
        return true;
    case kOptFuseDepth:
    case kOptLdsBudget: {
        if (value < 0 || (key == kOptLdsBudget && value > 0 && value < 4096)) {
            err = key == kOptFuseDepth ? "fusion depth must be >= 0" : "LDS budget must be 0 or >= 4096 bytes";
            return false;
        }
        if (key == kOptFuseDepth) {
            if (value == e.fuseDepth) return true;
            e.fuseDepth = value;
        } else {
            const size_t b = value == 0 ? e.ldsDevice : std::min<size_t>(e.ldsDevice, (size_t)value);
            if (b == e.ldsMax) return true;
            e.ldsMax = b;
        }
        if (e.M > 0) {
            const std::vector<int8_t> ops4 = e.ops4;
            const std::vector<int> order = e.order;
            const std::vector<double> scale = e.scale;
            if (!set_operators(ops4, order, scale, e.P, err)) return false;
        }
        e.bnvec = 0;
        return true;
    }
    case kOptLaneGroups:
    case kOptLaneGroupsFwd: {
        if (value < 1 || value > 3) {
            err = "lane groups must be 1, 2 or 3";
            return false;
        }
        int &dst = key == kOptLaneGroups ? e.maxGroups : e.maxGroupsF;
        if (value == dst) return true;
        dst = value;
        if (e.M > 0 && !e.cplx) {
            e.drop_graphs();
            if (!e.build_lane_segments(err)) return false;
        }
        e.bnvec = 0;
        return true;
    }
    case kOptSigmaHalves:
        if (value < 0 || value > 1) {
            err = "sigma halves option must be 0 or 1";
            return false;
        }
        if (value == e.sighOpt) return true;
        e.sighOpt = value;
        e.drop_graphs();
        return true;
    case kOptGemmSolutionBeta:
    case kOptGemmSolutionAlpha: {
        int &dst = e.gemmSol[key == kOptGemmSolutionAlpha ? 1 : 0];
        if (value == dst) return true;
        dst = value;
        e.drop_graphs();
        return true;
    }
    case kOptSigmaGrouped:
        if (value < 0 || value > 1) {
            err = "sigma grouped option must be 0 or 1";
            return false;
        }
        if (value == e.siggOpt) return true;
        e.siggOpt = value;
        e.drop_graphs();
        return true;
    case kOptSigma1Rows:
        if (value != 0 && value != 2 && value != 4 && value != 6 && value != 8) {
            err = "sigma rows per phase must be 0, 2, 4, 6 or 8";
            return false;
        }
        if (value != e.sig1Rows) {
            e.sig1Rows = value;
            e.drop_graphs();
        }
        return true;
    case kOptReachable:
        if (value < 0 || value > 1) {
            err = "reachable option must be 0 or 1";
            return false;
        }
        if (value == e.reachOpt) return true;
        e.reachOpt = value;
        e.supDirty = e.hessDirty = true;
        e.bnvec = 0;  // lists and stored states are renewed at the next upload
        e.drop_graphs();
        return true;
    case kOptCheckpoint:
        if (value < 0 || value > 1) {
            err = "checkpoint option must be 0 or 1";
            return false;
        }
        if (value == e.ckOpt) return true;
        e.ckOpt = value;
        e.bnvec = 0;  // decided at the next upload
        e.drop_graphs();
        return true;
    case kOptWindow:
    case kOptWindowWaves:
        if (key == kOptWindow ? (value != 0 && value != -1 && (value < 2 || value > kWinMaxW))
                              : (value != 4 && value != 8)) {
            err = key == kOptWindow ? "window width must be -1, 0 or 2 to 8" : "window waves per block must be 4 or 8";
            return false;
        }
        if (value == (key == kOptWindow ? e.winW : e.winWaves)) return true;
        (key == kOptWindow ? e.winW : e.winWaves) = value;
        if (e.M > 0 && !e.cplx) {
            e.drop_graphs();
            if (!e.build_lane_segments(err)) return false;
        }
        e.bnvec = 0;  // stored states and lists are renewed at the next upload
        return true;
    case kOptSigmaExchange:
        if (value != -1 && value != 0) {
            err = "exchange-symmetry option must be -1 or 0";
            return false;
        }
        if (value != e.exchOpt) {
            // off at once; on from the next upload (which checks the batch)
            e.exchOpt = value;
            if (value == 0 && (e.bSym || e.bSymV)) {
                e.bSym = e.bSymV = false;
                e.drop_graphs();
            }
        }
        return true;
    case kOptWindowStaged:
        if (value < 0 || value > 1) {
            err = "staged window option must be 0 or 1";
            return false;
        }
        if ((value == 1) != e.winVs) {
            e.winVs = value == 1;
            e.drop_graphs();
        }
        return true;
    case kOptWindowHessian:
        if (value < 0 || value > 1) {
            err = "window Hessian option must be 0 or 1";
            return false;
        }
        e.winHess = value == 1;
        return true;
    case kOptWindowEnergy:
        if (value < 0 || value > 1) {
            err = "window energy option must be 0 or 1";
            return false;
        }
        if ((value == 1) != e.winFuseE) {
            e.winFuseE = value == 1;
            e.drop_graphs();
        }
        return true;
    case kOptWindowVector:
        if (value < -1 || value > 1) {
            err = "vector-major window option must be -1, 0 or 1";
            return false;
        }
        if (value != e.winVec) {
            e.winVec = value;
            e.drop_graphs();
        }
        return true;
    case kOptWindowPasses:
    case kOptWindowRegisters:
    case kOptHamKind:
    case kOptSegments:
    case kOptLaneSegments:
    case kOptCheckpointActive:
    case kOptSigmaHalvesAvailable:
    case kOptSigmaGroupedAvailable:
    case kOptSweepAmpForward:
    case kOptSweepAmpAdjointLambda:
    case kOptSweepAmpAdjointState:
        err = "read-only engine property";
        return false;
    default:
        err = "unknown engine option";
        return false;
    }
}

int Engine::option(int key) const
{
    const EngineImpl &e = *d_;
    switch (key) {
    case kOptSweep: return e.sweepKind;
    case kOptSweepTwoGroupAvailable: return e.sweepS ? 1 : 0;
    case kOptSigma1Rows: return e.sig1Rows;
    case kOptVectorSegments:
        return (int)(e.use_wv() ? e.wpass.size() : e.use_s() ? e.laneS.segs.size() : e.segs.size());
    case kOptLaneMin: return e.laneMin;
    case kOptGraph: return e.useGraph ? 1 : 0;
    case kOptFuseDepth: return e.fuseDepth;
    case kOptLdsBudget: return (int)e.ldsMax;
    case kOptHamKind: return e.hkind;
    case kOptSegments: return (int)e.segs.size();
    case kOptLaneGroups: return e.maxGroups;
    case kOptLaneGroupsFwd: return e.maxGroupsF;
    case kOptCheckpoint: return e.ckOpt;
    case kOptCheckpointActive: return e.ckOn ? 1 : 0;
    case kOptSigmaHalves: return e.sighOpt;
    case kOptSigmaHalvesAvailable: return e.sigh ? 1 : 0;
    case kOptSigmaGrouped: return e.siggOpt;
    case kOptSigmaGroupedAvailable: return e.sigg ? 1 : 0;
    case kOptGemmSolutionBeta: return e.gemmSol[0];
    case kOptGemmSolutionAlpha: return e.gemmSol[1];
    case kOptReachable: return e.reachOpt;
    case kOptSweepAmpForward: return (int)std::min<long long>(e.ampF, INT32_MAX);
    case kOptSweepAmpAdjointLambda: return (int)std::min<long long>(e.ampAL, INT32_MAX);
    case kOptSweepAmpAdjointState: return (int)std::min<long long>(e.ampAP, INT32_MAX);
    case kOptLaneSegments: return (int)e.lsegs.size();
    case kOptWindow: return e.winW;
    case kOptWindowWaves: return e.winWaves;
    case kOptWindowPasses: return e.winOn ? (int)e.wpass.size() : 0;
    case kOptWindowVector: return e.winVec;
    case kOptWindowEnergy: return e.winFuseE ? 1 : 0;
    case kOptWindowHessian: return e.winHess ? 1 : 0;
    case kOptWindowStaged: return e.use_vs() ? 1 : (e.winVs ? 2 : 0);
    case kOptSigmaExchange: return e.bSym || e.bSymV ? 1 : (e.exchOpt != 0 ? 2 : 0);
    case kOptWindowRegisters: return e.winOn && e.winReg ? (e.winTu ? 2 : 1) : 0;
    default: return -1;
    }
}

bool Engine::tune_sigma_gemms(std::string &report, std::string &err)
{
    EngineImpl &e = *d_;
    report.clear();
    if (!(e.hkind == 3 && e.sigv && e.bnvec >= 1 && e.use_lanes(e.bnvec) && e.blas)) return true;
    const int nch = (e.bnvec + kLanes - 1) / kLanes;
    const long long ny = (long long)nch * e.N * kLanes;
    double *x = e.bl_psi.p;
    double *y = e.bl_lam.p;
    const unsigned fgrid = (unsigned)std::min<long long>((ny + kBlock - 1) / kBlock, 8192);
    constexpr int kCmpBlocks = 1024;
    // every solution's output is checked against the default solution's on
    // a fixed input before it may be timed: a solution whose result differs
    // by more than kTol relative to max |y|, or is not finite, is rejected
    constexpr double kTol = 1e-10;
    double *yref = nullptr, *cmp = nullptr;
    HCHK(hipMalloc(&yref, ny * sizeof(double)));
    if (hipMalloc(&cmp, 2 * kCmpBlocks * sizeof(double)) != hipSuccess) {
        (void)hipFree(yref);
        err = "hipMalloc of the comparison buffer failed";
        return false;
    }
    std::vector<double> hcmp(2 * kCmpBlocks);
    hipEvent_t t0 = nullptr, t1 = nullptr;
    std::string rep = "{";
    bool ok = hipEventCreate(&t0) == hipSuccess && hipEventCreate(&t1) == hipSuccess;
    // y = V x with solution sol into a cleared y; false when the library refuses it
    auto run_once = [&](int which, rocblas_int sol) -> bool {
        hipLaunchKernelGGL(k_fill, dim3(fgrid), dim3(kBlock), 0, e.stream, y, ny, 0.0);
        return hipGetLastError() == hipSuccess && e.sigma_gemm(which, x, y, nch, sol) == rocblas_status_success;
    };
    if (ok) {
        hipLaunchKernelGGL(k_pattern, dim3(fgrid), dim3(kBlock), 0, e.stream, x, ny);
        ok = hipGetLastError() == hipSuccess;
    }
    for (int which = 0; which < 2 && ok; which++) {
        rocblas_int cnt = 0;
        std::vector<rocblas_int> list;
        if (e.sigma_gemm(which, x, y, nch, 0, true, nullptr, &cnt) == rocblas_status_success && cnt > 0) {
            list.resize(cnt);
            if (e.sigma_gemm(which, x, y, nch, 0, true, list.data(), &cnt) != rocblas_status_success) list.clear();
            list.resize(std::min<size_t>(list.size(), (size_t)std::max<rocblas_int>(cnt, 0)));
        }
        list.insert(list.begin(), 0);  // the library's default first
        // the default solution's output is the reference
        if (!run_once(which, 0)) {
            ok = false;
            break;
        }
        hipLaunchKernelGGL(k_copy, dim3(fgrid), dim3(kBlock), 0, e.stream, y, yref, ny);
        if (hipGetLastError() != hipSuccess) {
            ok = false;
            break;
        }
        int best = 0, rejected = 0;
        float bestMs = 1e30f, defMs = 0;
        for (rocblas_int sol : list) {
            if (sol != 0) {
                if (!run_once(which, sol)) continue;
                hipLaunchKernelGGL(k_absdiff_max, dim3(kCmpBlocks), dim3(kBlock), 0, e.stream, y, yref, ny, cmp);
                if (hipGetLastError() != hipSuccess ||
                    hipMemcpyAsync(hcmp.data(), cmp, hcmp.size() * sizeof(double), hipMemcpyDeviceToHost,
                                   e.stream) != hipSuccess ||
                    hipStreamSynchronize(e.stream) != hipSuccess) {
                    ok = false;
                    break;
                }
                double dmax = 0, rmax = 0;
                for (int b = 0; b < kCmpBlocks; b++) {
                    dmax = std::max(dmax, hcmp[2 * b]);
                    rmax = std::max(rmax, hcmp[2 * b + 1]);
                }
                if (!(dmax <= kTol * std::max(1.0, rmax))) {
                    rejected++;
                    continue;
                }
            }
            // one warm-up call, then the mean of three
            if (e.sigma_gemm(which, x, y, nch, sol) != rocblas_status_success) continue;
            if (hipEventRecord(t0, e.stream) != hipSuccess) { ok = false; break; }
            bool good = true;
            for (int r = 0; r < 3 && good; r++) good = e.sigma_gemm(which, x, y, nch, sol) == rocblas_status_success;
            if (hipEventRecord(t1, e.stream) != hipSuccess || hipEventSynchronize(t1) != hipSuccess) {
                ok = false;
                break;
            }
            float ms = 0;
            if (!good || hipEventElapsedTime(&ms, t0, t1) != hipSuccess) continue;
            ms /= 3;
            if (sol == 0) defMs = ms;
            if (ms < bestMs) {
                bestMs = ms;
                best = sol;
            }
        }
        e.gemmSol[which] = best;
        char buf[200];
        std::snprintf(buf, sizeof(buf),
                      "%s\"%s\": {\"solutions\": %d, \"rejected\": %d, \"default_ms\": %.4f, \"chosen\": %d, "
                      "\"ms\": %.4f}",
                      which ? ", " : "", which ? "alpha" : "beta", (int)list.size() - 1, rejected, defMs, best, bestMs);
        rep += buf;
    }
    (void)hipStreamSynchronize(e.stream);
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
    (void)hipFree(yref);
    (void)hipFree(cmp);
    if (!ok) {
        err = "timing the sigma DGEMMs failed";
        return false;
    }
    e.drop_graphs();
    report = rep + "}";
    return true;
}

bool Engine::synchronize(std::string &err)
{
    HCHK(hipStreamSynchronize(d_->stream));
    d_->collect_marks();
    return true;
}

void Engine::set_timing(bool on)
{
    EngineImpl &e = *d_;
    e.timingOn = on;
    for (int c = 0; c < EngineImpl::kClasses; c++) {
        e.tms[c] = 0;
        e.tcnt[c] = 0;
    }
}

bool Engine::timing(int cls, double *ms, long long *launches, std::string &err)
{
    if (cls < 0 || cls >= EngineImpl::kClasses) {
        err = "unknown kernel class";
        return false;
    }
    if (!synchronize(err)) return false;
    *ms = d_->tms[cls];
    *launches = d_->tcnt[cls];
    return true;
}

bool Engine::batch(int nvec, const double *angles, double *E, double *deriv, std::string &err)
{
    return batch_upload(nvec, angles, err) && batch_run(deriv != nullptr, err) && batch_download(E, deriv, err);
}

// Per-parameter Hessian, d2E = 2<psi|H|d2 psi> + 2<d psi|H|d psi>.
//   1. psi = U psi_0, lambda = H psi, adjoint sweep: brackets give the
//      gradient and the sweep ends at lambda_0 = U^T H psi.
//   2. One forward sweep carries psi (from psi_0), lambda (from lambda_0) and
//      the per-parameter tangent accumulators B_mu; at operator k, after the
//      rotation, B_{order_k} += scale_k G_k psi, and
//      X[mu][k] = <lambda|G_k|B_mu>, D[k] = <lambda|G_k G_k|psi>.
//      At the end B_mu = d psi / d theta_mu.
//   3. THT = 2 B (H B)^T;  Y[mu][nu] = sum_{k in nu} scale_k X[mu][k];
//      Hessian = THT + 2 (Y + Y^T) - 2 diag(sum_{k in mu} scale_k^2 D[k]).
// This is the reference's quantity (src/fusedevolve.cpp:2437-2780) obtained
// with forward-propagated tangents instead of one adjoint sweep per angle.
bool Engine::hessian(const double *angles, double *hess, double *grad, double *E, std::string &err, double *metric)
{
    return second_derivatives(angles, nullptr, hess, grad, E, metric, err);
}

bool Engine::hessian_proj(const double *angles, const double *proj, double *hess, double *grad, double *overlap,
                          std::string &err)
{
    return second_derivatives(angles, proj, hess, grad, overlap, nullptr, err);
}

// Second derivatives by one forward sweep of tangents.  proj == nullptr:
// E = <psi|H|psi> (lambda_0 = U^T H psi, factor 2, plus the <dpsi|H|dpsi>
// term); otherwise f = <proj|psi> (lambda_0 = U^T proj, factor 1, no
// such term; src/fusedevolve.cpp:2892-3183).  metric (optional, energy
// only): g[mu][nu] = <dpsi/dtheta_mu|dpsi/dtheta_nu>.
// tangent count from which the Hessian's H B runs through the lane-layout sigma
constexpr int kHessLaneMin = 16;

bool Engine::second_derivatives(const double *angles, const double *proj, double *hess, double *grad, double *val,
                                double *metric, std::string &err)
{
    EngineImpl &e = *d_;
    if (!e.real_only("the Hessian", err)) return false;
    if (!e.haveStart || e.M == 0 || (!proj && e.hkind == 0)) {
        err = "engine not set up";
        return false;
    }
    const int M = e.M, P = e.P;
    const long long N = e.N;
    // 1. the first derivatives; afterwards d_lam = lambda_0 (the adjoint
    // state at the start) and d_phi ~ psi_0
    std::vector<double> deriv(M);
    double E0 = 0;
    if (proj) {
        if (!overlap_gradient(angles, proj, &E0, deriv.data(), err)) return false;
    } else {
        if (!evolve(angles, err)) return false;
        if (!gradient(angles, &E0, deriv.data(), err)) return false;
    }
    const double fac = proj ? 1.0 : 2.0;
    // parameters ordered by their first operator, so the active set at
    // operator k is a prefix
    std::vector<int> first(P, M);
    for (int k = 0; k < M; k++) first[e.order[k]] = std::min(first[e.order[k]], k);
    std::vector<int> perm(P);  // perm position -> parameter
    std::iota(perm.begin(), perm.end(), 0);
    std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return first[a] < first[b]; });
    std::vector<int> pos(P);
    for (int i = 0; i < P; i++) pos[perm[i]] = i;
    const int G = (int)std::min<long long>((N + kBlock - 1) / kBlock, 512);
    const size_t PN = (size_t)P * N;
    if (!e.hz_B.ensure(PN, err) || !e.hz_HB.ensure(PN, err) || !e.hz_psi2.ensure(2 * N, err) ||
        !e.hz_lam2.ensure(2 * N, err) || !e.hz_xpart.ensure((size_t)M * P * G, err) ||
        !e.hz_dpart.ensure((size_t)M * G, err) ||
        !e.hz_epart.ensure((size_t)P * e.eslots_alloc(), err))
        return false;
    auto fill = [&](double *p, size_t cnt) {
        hipLaunchKernelGGL(k_fill, dim3((unsigned)std::min<size_t>((cnt + kBlock - 1) / kBlock, 8192)), dim3(kBlock), 0,
                           e.stream, p, (long long)cnt, 0.0);
    };
    // 2. the tangents: in lane layout over two-group segments (one launch
    // per segment for every tangent, psi and lambda advanced by the
    // vector-major two-group kernel), else one launch per operator over the
    // tangents active so far
    // window passes (TUPS units in four-orbital register tiles) or two-group segments
    const bool tanWin = e.mode == kSz && P >= kHessLaneMin && e.winOn && e.winReg && e.winTu && e.winHess;
    const bool tanLanes = tanWin || (e.mode == kSz && e.laneSAllTwo && P >= kHessLaneMin);
    const int nch = (P + kLanes - 1) / kLanes;
    const int TS = std::max(1, e.numCU * 4);  // partial rows per operator of the lane tangent steps
    const size_t L = (size_t)nch * N * kLanes;
    if (tanLanes) {
        if (!e.hz_lx.ensure(L, err) || !e.hz_xpart.ensure((size_t)nch * M * TS * kLanes, err) ||
            !e.hz_dpart.ensure((size_t)M * TS, err) || !e.hz_inj.ensure(M, err) || !e.hz_scl.ensure(M, err))
            return false;
        std::vector<int> inj(M);
        for (int k = 0; k < M; k++) inj[k] = pos[e.order[k]];
        HCHK(hipMemcpy(e.hz_inj.p, inj.data(), sizeof(int) * M, hipMemcpyHostToDevice));
        HCHK(hipMemcpy(e.hz_scl.p, e.scale.data(), sizeof(double) * M, hipMemcpyHostToDevice));
        fill(e.hz_lx.p, L);
        fill(e.hz_xpart.p, (size_t)nch * M * TS * kLanes);
        fill(e.hz_dpart.p, (size_t)M * TS);
        LCHK();
        double *psi = e.hz_psi2.p, *lam = e.hz_lam2.p;
        HCHK(hipMemcpyAsync(psi, e.d_start.p, 8 * N, hipMemcpyDeviceToDevice, e.stream));
        HCHK(hipMemcpyAsync(lam, e.d_lam.p, 8 * N, hipMemcpyDeviceToDevice, e.stream));
        // window passes: every tile of a pass, psi and lambda read from one
        // half of their buffers and written by chunk 0 to the other
        double *psiB = psi + N, *lamB = lam + N;
        if (tanWin && !e.hessian_tiles(err)) return false;
        for (size_t p = 0; tanWin && p < e.wpass.size(); p++) {
            const auto &wp = e.wpass[p];
            WinArgs a = e.winargs(wp);
            a.tiles = e.d_wall.p + wp.allOff;
            a.ntiles = (int)wp.allTiles.size();
            const size_t lds = (size_t)2 * wp.nkp * (kLanes + 1) * sizeof(double);
            hipLaunchKernelGGL(k_wint, dim3((unsigned)std::min(wp.gridT, TS), nch), dim3(2 * kLanes), lds, e.stream, a,
                               psi, lam, psiB, lamB, e.hz_lx.p, (long long)N * kLanes, (const double2 *)e.d_cs.p, M,
                               e.hz_inj.p, e.hz_scl.p, e.hz_xpart.p, e.hz_dpart.p, TS, wp.nReach);
            LCHK();
            std::swap(psi, psiB);
            std::swap(lam, lamB);
        }
        for (const Seg &sg : e.laneS.segs) {
            if (tanWin) break;
            Orb2Args oa = e.orb2args(e.laneS, sg, false);
            oa.list = oa.llist = nullptr;
            TanArgs ta;
            ta.psi = psi;
            ta.lam = lam;
            ta.inj = e.hz_inj.p;
            ta.scale = e.hz_scl.p;
            ta.xpart = e.hz_xpart.p;
            ta.dpart = e.hz_dpart.p;
            ta.slots = TS;
            const int grid = (int)std::max<long long>(1, std::min<long long>((sg.orbits + 3) / 4, TS));
            hipLaunchKernelGGL(k_orb2t, dim3(grid, nch), dim3(kBlock), 0, e.stream, oa, ta, e.hz_lx.p,
                               (long long)N * kLanes, (const double2 *)e.d_cs.p, M);
            LCHK();
            hipLaunchKernelGGL((k_orb2s<false>), dim3(e.s_grid(sg), 1), dim3(kBlock), 0, e.stream, oa, psi,
                               (double *)nullptr, N, (const double2 *)e.d_cs.p, M, (double *)nullptr, e.slots);
            LCHK();
            hipLaunchKernelGGL((k_orb2s<false>), dim3(e.s_grid(sg), 1), dim3(kBlock), 0, e.stream, oa, lam,
                               (double *)nullptr, N, (const double2 *)e.d_cs.p, M, (double *)nullptr, e.slots);
            LCHK();
        }
    } else {
        fill(e.hz_B.p, PN);
        fill(e.hz_xpart.p, (size_t)M * P * G);
        LCHK();
        double *psiA = e.hz_psi2.p, *psiB = e.hz_psi2.p + N;
        double *lamA = e.hz_lam2.p, *lamB = e.hz_lam2.p + N;
        HCHK(hipMemcpyAsync(psiA, e.d_start.p, 8 * N, hipMemcpyDeviceToDevice, e.stream));
        HCHK(hipMemcpyAsync(lamA, e.d_lam.p, 8 * N, hipMemcpyDeviceToDevice, e.stream));
        const DevSector sec = e.sec();
        int nactive = 0;
        for (int k = 0; k < M; k++) {
            while (nactive < P && first[perm[nactive]] <= k) nactive++;
            const int slices = (nactive + 3) / 4;
            hipLaunchKernelGGL((k_tangent<4>), dim3(G, std::max(slices, 1)), dim3(kBlock), 0, e.stream, psiA, psiB,
                               lamA, lamB, e.hz_B.p, N, nactive, N, sec, e.d_ops.p, k, e.d_cs.p, pos[e.order[k]],
                               e.scale[k], e.hz_xpart.p + (size_t)k * P * G, G, e.hz_dpart.p + (size_t)k * G);
            LCHK();
            std::swap(psiA, psiB);
            std::swap(lamA, lamB);
        }
    }
    // 3. (energy) H B and the Gram matrix 2 B (H B)^T; the metric B B^T;
    // reductions
    const int chunks = 64;
    const int tiles = (P + 15) / 16;
    const bool withT = !proj, withG = metric && !proj;
    if (!e.hz_out.ensure((size_t)M * P + M + 2 * (size_t)P * P, err)) return false;
    double *outX = e.hz_out.p, *outD = outX + (size_t)M * P, *outT = outD + M, *outG = outT + (size_t)P * P;
    if (withT || withG) {
        if (!e.hz_gram.ensure((size_t)chunks * P * P, err)) return false;
    }
    const dim3 tgrid((unsigned)((N + kLanes - 1) / kLanes), nch);
    if (withT && e.hkind == 3 && e.sigv && P >= kHessLaneMin) {
        // H B through the lane-layout sigma (about ten times the vector-major
        // sigma's throughput per vector): B in lanes, sigma, H B back
        if (!e.hz_lx.ensure(L, err) || !e.hz_ly.ensure(L, err) || !e.hz_lep.ensure((size_t)512 * nch * kLanes, err))
            return false;
        if (!tanLanes) {
            hipLaunchKernelGGL(k_to_lanes, tgrid, dim3(kBlock), 0, e.stream, e.hz_B.p, e.hz_lx.p, N, P);
            LCHK();
        }
        double *zs = nullptr;
        if (e.tangent_symmetric(angles)) {
            if (!e.hz_z.ensure(L, err)) return false;
            zs = e.hz_z.p;
        }
        if (!e.ham_apply_lanes(e.hz_lx.p, e.hz_ly.p, nch, P, e.hz_epart.p, err, e.hz_lep.p, true, zs)) return false;
        hipLaunchKernelGGL(k_from_lanes, tgrid, dim3(kBlock), 0, e.stream, e.hz_ly.p, e.hz_HB.p, N, P);
        LCHK();
    }
    if (tanLanes && (withT || withG)) {
        // the tangents vector-major for the Gram matrices
        hipLaunchKernelGGL(k_from_lanes, tgrid, dim3(kBlock), 0, e.stream, e.hz_lx.p, e.hz_B.p, N, P);
        LCHK();
    }
    if (withT && !(e.hkind == 3 && e.sigv && P >= kHessLaneMin)) {
        if (!e.ham_apply(e.hz_B.p, e.hz_HB.p, P, e.hz_epart.p, err)) return false;
    }
    // Gram matrices out[a*P + b] = fac * sum_i X[a][i] Y[b][i]: one FP64 GEMM
    // where the engine has a rocBLAS handle (column-major view: out = fac *
    // Y^T X over vectors of stride N), else the tiled kernel
    auto gram = [&](const double *X, const double *Y, double *out, double f) -> bool {
        if (e.blas) {
            const double zero = 0.0;
            if (rocblas_dgemm(e.blas, rocblas_operation_transpose, rocblas_operation_none, P, P, (rocblas_int)N, &f, Y,
                              (rocblas_int)N, X, (rocblas_int)N, &zero, out, P) != rocblas_status_success) {
                err = "rocblas_dgemm (Gram matrix) failed";
                return false;
            }
            return true;
        }
        hipLaunchKernelGGL(k_gram, dim3(tiles, tiles, chunks), dim3(kBlock), 0, e.stream, X, Y, N, P, N, chunks,
                           e.hz_gram.p);
        LCHK();
        hipLaunchKernelGGL(k_reduce, dim3(P * P), dim3(kBlock), 0, e.stream, e.hz_gram.p, P * P, chunks, 1LL,
                           (long long)P * P, out, f);
        LCHK();
        return true;
    };
    if (withT && !gram(e.hz_B.p, e.hz_HB.p, outT, 2.0)) return false;
    if (withG && !gram(e.hz_B.p, e.hz_B.p, outG, 1.0)) return false;
    if (tanLanes) {
        // outX[pos * M + k]
        hipLaunchKernelGGL(k_reduce_lanes, dim3(nch * M), dim3(kBlock), 0, e.stream, e.hz_xpart.p, M, TS,
                           (const int *)nullptr, P, outX, 1.0);
        hipLaunchKernelGGL(k_reduce, dim3(M), dim3(kBlock), 0, e.stream, e.hz_dpart.p, M, TS, (long long)TS, 1LL,
                           outD, 1.0);
    } else {
        // outX[k * P + pos]
        hipLaunchKernelGGL(k_reduce, dim3(M * P), dim3(kBlock), 0, e.stream, e.hz_xpart.p, M * P, G, (long long)G,
                           1LL, outX, 1.0);
        hipLaunchKernelGGL(k_reduce, dim3(M), dim3(kBlock), 0, e.stream, e.hz_dpart.p, M, G, (long long)G, 1LL, outD,
                           1.0);
    }
    LCHK();
    std::vector<double> hout((size_t)M * P + M + 2 * (size_t)P * P, 0.0);
    const size_t nread = (size_t)M * P + M + (withG ? 2 : withT ? 1 : 0) * (size_t)P * P;
    HCHK(hipMemcpyAsync(hout.data(), e.hz_out.p, 8 * nread, hipMemcpyDeviceToHost, e.stream));
    HCHK(hipStreamSynchronize(e.stream));
    const double *X = hout.data(), *D = X + (size_t)M * P, *T = D + M, *Gm = T + (size_t)P * P;
    // assemble in parameter order
    std::vector<double> Y((size_t)P * P, 0.0), Dg(P, 0.0);
    for (int k = 0; k < M; k++) {
        const int nu = e.order[k];
        for (int mu = 0; mu < P; mu++)
            Y[(size_t)mu * P + nu] += e.scale[k] * X[tanLanes ? (size_t)pos[mu] * M + k : (size_t)k * P + pos[mu]];
        Dg[nu] += e.scale[k] * e.scale[k] * D[k];
    }
    for (int i = 0; i < P; i++)
        for (int j = 0; j < P; j++) {
            const int a = std::min(pos[i], pos[j]), b = std::max(pos[i], pos[j]);
            double v = (withT ? T[(size_t)a * P + b] : 0.0) + fac * (Y[(size_t)i * P + j] + Y[(size_t)j * P + i]);
            if (i == j) v -= fac * Dg[i];
            hess[(size_t)i * P + j] = v;
            if (withG) metric[(size_t)i * P + j] = Gm[(size_t)a * P + b];
        }
    if (grad) {
        for (int p = 0; p < P; p++) grad[p] = 0;
        for (int k = 0; k < M; k++) grad[e.order[k]] += e.scale[k] * deriv[k];
    }
    if (val) *val = E0;
    return true;
}

} // namespace ae
