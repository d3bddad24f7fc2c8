// Shared definitions of the MI355X statevector engine.
#pragma once
#include <cstdint>
#include <cstdio>
#include <string>

namespace ae {

// Sector kinds, chosen exactly as the reference's manager chooses its
// compressor (src/AnsatzManager.cpp:88-100): particle number and S_z fixed,
// particle number only, or the full Fock space.
enum SectorMode : int { kFull = 0, kNumber = 1, kSz = 2 };

// One excitation operator on spin-orbital bit strings, reduced to the three
// masks the pair rule needs (src/fusedevolve.cpp:27-145):
//   b is paired with b' = b ^ cre ^ ann when (b & (cre|ann)) == ann (the
//   "annihilate" side, phase +1) or == cre (phase -1); the phase is further
//   multiplied by (-1)^popc(b & mask) and by `flip` (-1 per out-of-order
//   index pair, a0 > a1 and a2 > a3).
struct DevOp {
    uint32_t cre;
    uint32_t ann;
    uint32_t mask;
    int32_t flip;
};

// Builds the DevOp of a validated operator (a: zero-based, -1 = absent).
// Returns false for operators the reference cannot evolve: anything that
// is not a single a^+_a a_b (a != b) or a double with four distinct
// indices, or that addresses a spin-orbital >= nq.  With numberOps (complex
// amplitudes) number operators are accepted as well: a^+_p a_p and doubles
// whose created pair is the annihilated pair (src/fusedevolve.cpp:82-95;
// other repeated indices stay unsupported, src/fusedevolve.cpp:1806).  Their
// DevOp has cre == ann == the orbitals that must be occupied, mask 0, and
// flip the ordering phase (-1: the operator never acts).
inline bool make_devop(const int8_t a[4], int nq, DevOp &o, std::string &why, bool numberOps = false)
{
    auto inr = [nq](int x) { return x >= 0 && x < nq; };
    if (a[2] < 0 && a[3] < 0) {
        if (!inr(a[0]) || !inr(a[1])) { why = "single excitation index out of range"; return false; }
        if (a[0] == a[1] && numberOps) {
            o.cre = o.ann = 1u << a[0];
            o.mask = 0;
            o.flip = 1;
            return true;
        }
        if (a[0] == a[1]) { why = "number operators need complex amplitudes"; return false; }
        o.cre = 1u << a[0];
        o.ann = 1u << a[1];
        uint32_t m = ((1u << a[0]) - 1) ^ ((1u << a[1]) - 1);
        o.mask = m & ~(o.cre | o.ann);
        o.flip = 1;
        return true;
    }
    if (a[2] < 0 || a[3] < 0) { why = "operator is neither a single nor a double excitation"; return false; }
    for (int i = 0; i < 4; i++)
        if (!inr(a[i])) { why = "double excitation index out of range"; return false; }
    if (numberOps && a[0] != a[1] && a[2] != a[3] &&
        ((1u << a[0]) | (1u << a[1])) == ((1u << a[2]) | (1u << a[3]))) {
        o.cre = o.ann = (1u << a[0]) | (1u << a[1]);
        o.mask = 0;
        o.flip = ((a[0] > a[1]) ? -1 : 1) * ((a[2] > a[3]) ? -1 : 1);
        return true;
    }
    for (int i = 0; i < 4; i++)
        for (int j = i + 1; j < 4; j++)
            if (a[i] == a[j]) { why = "double excitation with a repeated index is not a supported operator"; return false; }
    o.cre = (1u << a[0]) | (1u << a[1]);
    o.ann = (1u << a[2]) | (1u << a[3]);
    uint32_t m = ((1u << a[0]) - 1) ^ ((1u << a[1]) - 1) ^ ((1u << a[2]) - 1) ^ ((1u << a[3]) - 1);
    o.mask = m & ~(o.cre | o.ann);
    o.flip = ((a[0] > a[1]) ? -1 : 1) * ((a[2] > a[3]) ? -1 : 1);
    return true;
}

// Host form of the pair rule: 0 if the operator does not act on b, else the
// phase and the partner in b2.
inline int host_pair_sign(const DevOp &o, uint32_t b, uint32_t &b2)
{
    const uint32_t m = b & (o.cre | o.ann);
    int s;
    if (m == o.ann) s = 1;
    else if (m == o.cre) s = -1;
    else return 0;
    b2 = b ^ o.cre ^ o.ann;
    if (__builtin_popcount(b & o.mask) & 1) s = -s;
    return s * o.flip;
}

void log_msg(const char *fmt, ...);

} // namespace ae
