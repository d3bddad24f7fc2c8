// Autotuner of the batched E+grad sequence.
//
// Candidates are execution choices that leave the mathematics unchanged:
// the sweep kind (register orbits, LDS-staged orbits, LDS tiles only), the
// fusion depth (most operators per pass), the batch layout (lane layout
// or one vector per grid slice) and, in the lane layout, orbit segments or
// window passes.  Each candidate runs the caller's angle
// batch once to warm up (and capture its hipGraph), then `reps` timed
// replays; the median wall time decides among the candidates whose results
// agree.  Every candidate's energies and per-operator derivatives are
// compared with the first candidate's: a candidate whose largest deviation,
// relative to max(1, |first value|), exceeds kAgreeTol, or whose results are
// not finite, is reported as rejected and never chosen or recorded.  The
// first candidate must itself give finite results.  A choice can be
// recorded per problem shape in a text file of JSON lines and re-applied
// without measuring.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>

#include "manager.hpp"

namespace ae {

void log_msg(const char *fmt, ...);

namespace {

// largest accepted deviation of a candidate's results from the first
// candidate's (relative to max(1, |value|)); execution choices change only
// the summation order, which moves results by ~1e-15
constexpr double kAgreeTol = 1e-10;

// largest relative deviation of (E, D) from (E0, D0); +inf when any value is
// not finite
double deviation(const std::vector<double> &E, const std::vector<double> &E0, const std::vector<double> &D,
                 const std::vector<double> &D0)
{
    double dev = 0;
    auto acc = [&](double x, double x0) {
        const double d = std::fabs(x - x0) / std::max(1.0, std::fabs(x0));
        dev = std::isfinite(d) ? std::max(dev, d) : INFINITY;
    };
    for (size_t i = 0; i < E.size(); i++) acc(E[i], E0[i]);
    for (size_t i = 0; i < D.size(); i++) acc(D[i], D0[i]);
    return dev;
}

bool all_finite(const std::vector<double> &a)
{
    for (double x : a)
        if (!std::isfinite(x)) return false;
    return true;
}

struct Candidate {
    int sweep, depth, laneMin, groups, groupsFwd, ckpt, halves, grouped;
    int window = 0;            // kOptWindow: 0 orbit segments, -1 automatic window passes
    int gemmB = 0, gemmA = 0;  // rocBLAS solutions of the sigma DGEMMs (0: default; tuned after the choice)
};

std::string shape_key(const Engine &eng, int nvec, int nops)
{
    std::ostringstream k;
    k << "N=" << eng.dim() << " M=" << nops << " ham=" << eng.ham_kind() << " nvec=" << nvec;
    return k.str();
}

std::string cand_json(const Candidate &c)
{
    std::ostringstream o;
    o << "{\"sweep\": " << c.sweep << ", \"depth\": " << c.depth << ", \"lane_min\": " << c.laneMin
      << ", \"groups\": " << c.groups << ", \"groups_fwd\": " << c.groupsFwd << ", \"checkpoint\": " << c.ckpt
      << ", \"sigma_halves\": " << c.halves << ", \"sigma_grouped\": " << c.grouped << ", \"window\": " << c.window
      << ", \"gemm_beta\": " << c.gemmB << ", \"gemm_alpha\": " << c.gemmA << "}";
    return o.str();
}

// Last record of `key` in the cache file, if any.
bool lookup(const std::string &file, const std::string &key, Candidate &c)
{
    std::ifstream in(file);
    if (!in) return false;
    std::string line;
    bool found = false;
    const std::string tag = "\"key\": \"" + key + "\"";
    while (std::getline(in, line)) {
        if (line.find(tag) == std::string::npos) continue;
        Candidate t{};
        if (std::sscanf(line.c_str() + line.find("\"choice\""),
                        "\"choice\": {\"sweep\": %d, \"depth\": %d, \"lane_min\": %d, \"groups\": %d, "
                        "\"groups_fwd\": %d, \"checkpoint\": %d, \"sigma_halves\": %d, \"sigma_grouped\": %d, "
                        "\"window\": %d, \"gemm_beta\": %d, \"gemm_alpha\": %d}",
                        &t.sweep, &t.depth, &t.laneMin, &t.groups, &t.groupsFwd, &t.ckpt, &t.halves, &t.grouped,
                        &t.window, &t.gemmB, &t.gemmA) == 11) {
            c = t;
            found = true;
        }
    }
    return found;
}

bool apply(Engine &eng, const Candidate &c, std::string &err)
{
    return eng.set_option(kOptSweep, c.sweep, err) && eng.set_option(kOptFuseDepth, c.depth, err) &&
           eng.set_option(kOptLaneGroups, c.groups, err) && eng.set_option(kOptLaneGroupsFwd, c.groupsFwd, err) &&
           eng.set_option(kOptLaneMin, c.laneMin, err) && eng.set_option(kOptCheckpoint, c.ckpt, err) &&
           eng.set_option(kOptSigmaHalves, c.halves, err) && eng.set_option(kOptSigmaGrouped, c.grouped, err) &&
           eng.set_option(kOptWindow, c.window, err) &&
           eng.set_option(kOptGemmSolutionBeta, c.gemmB, err) && eng.set_option(kOptGemmSolutionAlpha, c.gemmA, err) &&
           eng.set_option(kOptGraph, 1, err);
}

} // namespace

bool Manager::autotune(int nvec, int nangles, const double *angles, const char *cacheFile, int reps,
                       std::string &report)
{
    if (!ensureBuilt() || !realOnly("The autotuner")) return false;
    std::vector<double> dec;
    if (!decompressBatch(nvec, nangles, angles, dec)) return false;
    Engine &eng = *eng_;
    const int M = (int)ops_.size();
    const std::string key = shape_key(eng, nvec, M);
    std::string err;
    Candidate best{};
    if (cacheFile && *cacheFile && lookup(cacheFile, key, best)) {
        if (!apply(eng, best, err)) {
            log_msg("Autotune: applying the recorded choice failed: %s", err.c_str());
            return false;
        }
        report = "{\"key\": \"" + key + "\", \"source\": \"cache\", \"choice\": " + cand_json(best) + "}";
        return true;
    }
    reps = std::max(1, reps);

    std::vector<Candidate> cands;
    const bool sz = eng.mode() == kSz;
    const int sweeps[4] = {kSweepOrbit, kSweepOrbitLds, kSweepTile, kSweepOrbit2};
    // the lane layout segments kSweepOrbit and kSweepOrbit2 alike; its
    // candidates keep the engine's choice between the two for the
    // vector-major (single-vector) sweeps
    const int laneOrbit = eng.option(kOptSweep) == kSweepOrbit2 ? kSweepOrbit2 : kSweepOrbit;
    const int depths[3] = {0, 2, 1};
    const int gr = eng.option(kOptSigmaGroupedAvailable) == 1 ? 1 : 0;  // grouped sigma where it applies
    for (int lanes = 0; lanes < (nvec > 1 ? 2 : 1); lanes++)
        for (int s0 : sweeps) {
            if (!sz && s0 != kSweepOrbit) continue;          // other sectors sweep one operator per pass
            if (lanes && (s0 == kSweepOrbitLds || s0 == kSweepOrbit2)) continue;  // orbits in registers
            if (!lanes && s0 == kSweepOrbit2 && eng.option(kOptSweepTwoGroupAvailable) != 1) continue;
            const int s = lanes && s0 == kSweepOrbit ? laneOrbit : s0;
            for (int d : depths) {
                if (!sz && d != 0) continue;
                if (s0 == kSweepOrbit2 && d != 0) continue;
                cands.push_back({s, d, lanes ? 1 : 0, 1, 1, 0, 0, gr});
                // the lane layout can also fuse two disjoint groups per pass
                // (forward only, or both sweeps), keep the forward states for
                // the gradient, and use the other sigma variants where they apply
                if (lanes && sz && s0 == kSweepOrbit && d == 0) {
                    cands.push_back({s, d, 1, 1, 2, 0, 0, gr});
                    cands.push_back({s, d, 1, 2, 2, 0, 0, gr});
                    cands.push_back({s, d, 1, 1, 1, 1, 0, gr});
                    cands.push_back({s, d, 1, 2, 2, 1, 0, gr});
                    cands.push_back({s, d, 1, 3, 3, 1, 0, gr});
                    if (gr) cands.push_back({s, d, 1, 2, 2, 1, 0, 0});
                    if (eng.option(kOptSigmaHalvesAvailable) == 1) cands.push_back({s, d, 1, 2, 2, 1, 1, 0});
                    // window passes where their tiles fit in registers
                    // (the engine's automatic choice), with and without
                    // the stored states
                    cands.push_back({s, d, 1, 2, 2, 1, 0, gr, -1});
                    cands.push_back({s, d, 1, 2, 2, 0, 0, gr, -1});
                }
            }
        }

    Candidate saved{eng.option(kOptSweep),         eng.option(kOptFuseDepth),     eng.option(kOptLaneMin),
                    eng.option(kOptLaneGroups),    eng.option(kOptLaneGroupsFwd), eng.option(kOptCheckpoint),
                    eng.option(kOptSigmaHalves),   eng.option(kOptSigmaGrouped)};
    saved.window = eng.option(kOptWindow);
    saved.gemmB = eng.option(kOptGemmSolutionBeta);
    saved.gemmA = eng.option(kOptGemmSolutionAlpha);
    std::vector<double> E0(nvec), D0((size_t)nvec * M), E(nvec), D((size_t)nvec * M);
    std::ostringstream rep;
    rep << "{\"key\": \"" << key << "\", \"source\": \"measured\", \"reps\": " << reps << ", \"candidates\": [";
    double bestMs = 1e300, worstDev = 0;
    int nrejected = 0;
    bool haveBest = false;
    for (size_t ci = 0; ci < cands.size(); ci++) {
        const Candidate &c = cands[ci];
        if (!apply(eng, c, err) || !eng.batch_upload(nvec, dec.data(), err) || !eng.batch_run(true, err) ||
            !eng.batch_download(E.data(), D.data(), err)) {
            log_msg("Autotune candidate %s failed: %s", cand_json(c).c_str(), err.c_str());
            std::string e2;
            apply(eng, saved, e2);
            return false;
        }
        std::vector<double> ms;
        for (int r = 0; r < reps; r++) {
            const auto t0 = std::chrono::steady_clock::now();
            if (!eng.batch_run(true, err) || !eng.synchronize(err)) {
                log_msg("Autotune run failed: %s", err.c_str());
                return false;
            }
            ms.push_back(std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
        }
        if (!eng.batch_download(E.data(), D.data(), err)) return false;
        std::sort(ms.begin(), ms.end());
        const double med = ms[ms.size() / 2];
        double dev = 0;
        if (ci == 0) {
            if (!all_finite(E) || !all_finite(D)) {
                log_msg("Autotune: the first candidate %s gave non-finite results", cand_json(c).c_str());
                std::string e2;
                apply(eng, saved, e2);
                return false;
            }
            E0 = E;
            D0 = D;
        } else {
            dev = deviation(E, E0, D, D0);
        }
        const bool rejected = !(dev <= kAgreeTol);
        if (rejected) {
            nrejected++;
            log_msg("Autotune: candidate %s rejected, results deviate by %g", cand_json(c).c_str(), dev);
        } else {
            worstDev = std::max(worstDev, dev);
        }
        rep << (ci ? ", " : "") << "{\"choice\": " << cand_json(c) << ", \"ms\": " << med << ", \"max_dev\": "
            << (std::isfinite(dev) ? dev : 1e308) << ", \"rejected\": " << (rejected ? "true" : "false") << "}";
        if (!rejected && med < bestMs) {
            bestMs = med;
            best = c;
            haveBest = true;
        }
    }
    if (!haveBest) {  // cannot happen: the first candidate agrees with itself
        std::string e2;
        apply(eng, saved, e2);
        return false;
    }
    if (!apply(eng, best, err)) {
        log_msg("Autotune: applying the choice failed: %s", err.c_str());
        return false;
    }
    // rocBLAS solutions of the lane sigma's DGEMMs for the chosen layout
    // (every candidate above ran the library's default ones)
    std::string gemmRep;
    if (!eng.batch_upload(nvec, dec.data(), err) || !eng.tune_sigma_gemms(gemmRep, err)) {
        log_msg("Autotune: tuning the sigma DGEMMs failed: %s", err.c_str());
        return false;
    }
    best.gemmB = eng.option(kOptGemmSolutionBeta);
    best.gemmA = eng.option(kOptGemmSolutionAlpha);
    rep << "], \"max_dev\": " << worstDev << ", \"rejected\": " << nrejected << ", \"tolerance\": " << kAgreeTol
        << ", \"choice\": " << cand_json(best) << ", \"ms\": " << bestMs;
    if (!gemmRep.empty()) rep << ", \"sigma_gemms\": " << gemmRep;
    rep << "}";
    report = rep.str();
    if (cacheFile && *cacheFile) {
        std::ofstream out(cacheFile, std::ios::app);
        if (out)
            out << "{\"key\": \"" << key << "\", \"choice\": " << cand_json(best) << ", \"ms\": " << bestMs
                << ", \"max_dev\": " << worstDev << "}\n";
        else
            log_msg("Autotune: cannot append to %s", cacheFile);
    }
    return true;
}

bool Manager::setEngineOption(int key, int value)
{
    if (!ensureBuilt()) return false;
    std::string err;
    if (!eng_->set_option(key, value, err)) {
        log_msg("Engine option %d = %d rejected: %s", key, value, err.c_str());
        return false;
    }
    return true;
}

bool Manager::getEngineOption(int key, int &value)
{
    if (!ensureBuilt()) return false;
    value = eng_->option(key);
    // rocBLAS solution indices and the window choices (-1: automatic) may be
    // negative; every other option is >= 0 and -1 marks an unknown key
    if (key == kOptGemmSolutionBeta || key == kOptGemmSolutionAlpha || key == kOptWindow || key == kOptWindowVector)
        return true;
    return value >= 0;
}

} // namespace ae
