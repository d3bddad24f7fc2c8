// Command-line front end over the C ABI, with the argument words of the
// reference's cppAnsatzSynth (src/main.cpp:93-219):
//
//   ansatzsynth filepath <prefix> [optimise] [iterativeoptimise]
//               [subspacediag] [writeproperties [NoHess] [NoHFPath]
//               [NoLowestEigenValue] [NOverlap N]] [benchmark] [onlyevolve]
//               [loadpaths N] [loadExtractedMin] [help]
//
// It reads the run files <prefix>_* (start state, operators, order file,
// Hamiltonian as text triplets or integrals, nuclear energy, parameter
// sets) and, like the reference, takes path 0 as all angles zero and path
// i >= 1 as the i-th parameter set (src/TUPSLoadingUtils.cpp:226-266).
// Without a mode it prints the energy of every path; optimise runs the
// Newton optimiser from path 1 (src/main.cpp:488-499); benchmark times
// energy + gradient evaluations at path 1 (src/main.cpp:396-425);
// onlyevolve evolves path 1 and prints the final state's norm
// (src/main.cpp:380-392); iterativeoptimise runs the iterative projected
// optimiser from path 1; subspacediag diagonalises H in the span of paths
// 1..9 (src/main.cpp:438-484); writeproperties prints the properties table
// of every path (src/main.cpp:521-529); loadExtractedMin takes the paths
// from <prefix>_extractedmin instead of the parameter file
// (src/TUPSLoadingUtils.cpp:298-360).  Energies are printed electronic and
// with the nuclear energy.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "ansatzsynth.h"

namespace {

struct Options {
    std::string filePath;
    bool optimise = false, benchmark = false, onlyEvolve = false, help = false, ok = true;
    bool iterative = false, subspace = false, properties = false, extractedMin = false;
    bool noHess = false, noHFPath = false, noLowest = false;
    int overlaps = 1;
    int pathsToLoad = -1;
    int benchLoops = 100;
};

void usage()
{
    std::printf(
        "ansatzsynth filepath <prefix> [optimise] [benchmark] [onlyevolve] [loadpaths N] [help]\n"
        "  filepath <prefix>  run files <prefix>_Initial.dat, _Operators.dat, _Order.dat,\n"
        "                     _Parameters.dat, _Nuclear_Energy.dat and _Ham_Coeff.dat + _Ham_Index.dat\n"
        "                     or _oneEInts.bin + _twoEInts.bin\n"
        "  (no mode)          energy of every path (0: all angles zero, i: parameter set i)\n"
        "  optimise           Newton optimiser from path 1\n"
        "  benchmark [K]      time K (default 100) energy + gradient evaluations at path 1\n"
        "  onlyevolve         evolve path 1, print the final state's norm\n"
        "  iterativeoptimise  iterative projected optimiser from path 1\n"
        "  subspacediag       eigenvalues of H in the span of paths 1..9\n"
        "  writeproperties    properties of every path: energies, Hessian and metric eigenvalue counts,\n"
        "                     gradient norm, overlaps with the lowest eigenvectors of H\n"
        "    NoHess             skip the Hessian and metric\n"
        "    NoHFPath           leave out path 0\n"
        "    NoLowestEigenValue no eigenvectors of H\n"
        "    NOverlap N         overlaps with the N lowest eigenvectors (default 1)\n"
        "  loadpaths N        keep the first N paths (path 0 included)\n"
        "  loadExtractedMin   paths from <prefix>_extractedmin (parameter sets, one value per parameter)\n");
}

Options parse(int argc, char **argv)
{
    Options o;
    for (int i = 1; i < argc; i++) {
        const char *a = argv[i];
        if (!std::strcmp(a, "optimise")) {
            o.optimise = true;
        } else if (!std::strcmp(a, "benchmark")) {
            o.benchmark = true;
            if (i + 1 < argc && std::atoi(argv[i + 1]) > 0) o.benchLoops = std::atoi(argv[++i]);
        } else if (!std::strcmp(a, "onlyevolve")) {
            o.onlyEvolve = true;
        } else if (!std::strcmp(a, "filepath") && i + 1 < argc) {
            o.filePath = argv[++i];
        } else if (!std::strcmp(a, "loadpaths") && i + 1 < argc) {
            o.pathsToLoad = std::atoi(argv[++i]);
        } else if (!std::strcmp(a, "iterativeoptimise")) {
            o.iterative = true;
        } else if (!std::strcmp(a, "subspacediag")) {
            o.subspace = true;
        } else if (!std::strcmp(a, "writeproperties")) {
            o.properties = true;
        } else if (!std::strcmp(a, "NoHess")) {
            o.noHess = true;
        } else if (!std::strcmp(a, "NoHFPath")) {
            o.noHFPath = true;
        } else if (!std::strcmp(a, "NoLowestEigenValue")) {
            o.noLowest = true;
        } else if (!std::strcmp(a, "NOverlap") && i + 1 < argc) {
            o.overlaps = std::atoi(argv[++i]);
            if (o.overlaps < 1) o.overlaps = 1;
        } else if (!std::strcmp(a, "loadExtractedMin")) {
            o.extractedMin = true;
        } else if (!std::strcmp(a, "help")) {
            o.help = true;
        } else {
            std::fprintf(stderr, "unknown argument: %s\n", a);
            o.ok = false;
        }
    }
    return o;
}

double read_nuclear(const std::string &prefix)
{
    double e = 0;
    if (FILE *f = std::fopen((prefix + "_Nuclear_Energy.dat").c_str(), "r")) {
        if (std::fscanf(f, "%lf", &e) != 1) e = 0;
        std::fclose(f);
    }
    return e;
}

// the qubit count: the length of the first bitstring of <prefix>_Initial.dat
int qubits_of(const std::string &prefix)
{
    int n = 0;
    if (FILE *f = std::fopen((prefix + "_Initial.dat").c_str(), "r")) {
        int c;
        while ((c = std::fgetc(f)) != EOF && (c == ' ' || c == '\t' || c == '\n' || c == '\r')) {
        }
        while (c == '0' || c == '1') {
            n++;
            c = std::fgetc(f);
        }
        std::fclose(f);
    }
    return n;
}

// Parameter sets of <prefix>_extractedmin: whitespace-separated numbers,
// nparams per set (src/TUPSLoadingUtils.cpp:298-360); an incomplete last
// set is dropped.
bool read_extracted_min(const std::string &prefix, int nparams, std::vector<std::vector<double>> &sets)
{
    FILE *f = std::fopen((prefix + "_extractedmin").c_str(), "r");
    if (!f) {
        std::fprintf(stderr, "cannot open %s_extractedmin\n", prefix.c_str());
        return false;
    }
    std::vector<double> vals;
    double v;
    while (std::fscanf(f, "%lf", &v) == 1) vals.push_back(v);
    std::fclose(f);
    sets.clear();
    for (size_t k = 0; k + nparams <= vals.size(); k += nparams) sets.emplace_back(vals.begin() + k, vals.begin() + k + nparams);
    if (vals.size() % (size_t)nparams)
        std::fprintf(stderr, "%s_extractedmin: %zu trailing values are not a whole set\n", prefix.c_str(),
                     vals.size() % (size_t)nparams);
    return true;
}

} // namespace

int main(int argc, char **argv)
{
    Options opt = parse(argc, argv);
    if (opt.help || argc <= 1) {
        usage();
        return opt.help ? 0 : 1;
    }
    if (!opt.ok) return 1;
    if (opt.filePath.empty()) {
        std::fprintf(stderr, "No file given\n");
        return 1;
    }
    void *ctx = init();
    int rc = loadRunFiles(opt.filePath.c_str(), 0, ctx);
    if (rc != 0) {
        std::fprintf(stderr, "loadRunFiles failed (%d)\n", rc);
        cleanup(&ctx);
        return 1;
    }
    long long dim = 0;
    int nops = 0, nparams = 0;
    if (getDimensions(&dim, &nops, &nparams, ctx) != 0) {
        std::fprintf(stderr, "construction failed\n");
        cleanup(&ctx);
        return 1;
    }
    const double nuc = read_nuclear(opt.filePath);
    std::printf("sector dimension %lld, operators %d, parameters %d, nuclear energy %.15f\n", dim, nops, nparams, nuc);

    // paths: 0 = zero angles, then the parameter sets
    std::vector<std::vector<double>> paths(1, std::vector<double>(nparams, 0.0));
    std::vector<double> recorded(1, NAN);
    int nsets = 0, np = 0;
    const std::string pfile = opt.filePath + "_Parameters.dat";
    if (readParameterSets(pfile.c_str(), 0, 0, &nsets, &np, nullptr, nullptr) == 0 && nsets > 0) {
        std::vector<double> E(nsets), P((size_t)nsets * np);
        if (readParameterSets(pfile.c_str(), nsets, np, &nsets, &np, E.data(), P.data()) == 0) {
            for (int s = 0; s < nsets; s++) {
                std::vector<double> p(nparams, 0.0);
                for (int i = 0; i < nparams && i < np; i++) p[i] = P[(size_t)s * np + i];
                paths.push_back(p);
                recorded.push_back(E[s]);
            }
        }
    }
    if (opt.extractedMin) {
        std::vector<std::vector<double>> sets;
        if (!read_extracted_min(opt.filePath, nparams, sets)) {
            cleanup(&ctx);
            return 1;
        }
        paths.resize(1);
        recorded.resize(1);
        for (auto &p : sets) {
            paths.push_back(p);
            recorded.push_back(NAN);
        }
    }
    if (opt.pathsToLoad > -1 && (size_t)opt.pathsToLoad < paths.size()) paths.resize(opt.pathsToLoad);
    const size_t first = paths.size() > 1 ? 1 : 0;

    if (opt.onlyEvolve) {
        // the final state is written over the full space of 2^qubits basis states
        const int nq = qubits_of(opt.filePath);
        if (nq < 1 || nq > 26) {
            std::fprintf(stderr, "onlyevolve: %d qubits (1 to 26 supported)\n", nq);
            cleanup(&ctx);
            return 1;
        }
        std::vector<double> st((size_t)1 << nq);
        if (getFinalState(nparams, paths[first].data(), 1 << nq, st.data(), ctx) != 0) {
            std::fprintf(stderr, "getFinalState failed\n");
            cleanup(&ctx);
            return 1;
        }
        double n2 = 0;
        for (double x : st) n2 += x * x;
        std::printf("path %zu: final state over 2^%d basis states, norm %.15f\n", first, nq, std::sqrt(n2));
        cleanup(&ctx);
        return 0;
    }
    if (opt.benchmark) {
        std::vector<double> g(nparams);
        double E = 0;
        getEnergyGradient(nparams, paths[first].data(), &E, g.data(), ctx);  // warm-up
        auto t0 = std::chrono::high_resolution_clock::now();
        for (int i = 0; i < opt.benchLoops; i++) getEnergyGradient(nparams, paths[first].data(), &E, g.data(), ctx);
        auto t1 = std::chrono::high_resolution_clock::now();
        const double us = std::chrono::duration<double, std::micro>(t1 - t0).count() / opt.benchLoops;
        std::printf("Benchmark Derivative Time per derivative (us) %.3f\n", us);
        t0 = std::chrono::high_resolution_clock::now();
        for (int i = 0; i < opt.benchLoops; i++) getEnergy(nparams, paths[first].data(), &E, ctx);
        t1 = std::chrono::high_resolution_clock::now();
        std::printf("Benchmark getEnergy (us) %.3f\n",
                    std::chrono::duration<double, std::micro>(t1 - t0).count() / opt.benchLoops);
        cleanup(&ctx);
        return 0;
    }
    for (size_t p = 0; p < paths.size(); p++) {
        double E = 0;
        if (getEnergy(nparams, paths[p].data(), &E, ctx) != 0) {
            std::fprintf(stderr, "getEnergy failed for path %zu\n", p);
            cleanup(&ctx);
            return 1;
        }
        std::printf("path %zu: energy %.15f (with nuclear %.15f)", p, E, E + nuc);
        if (!std::isnan(recorded[p])) std::printf(", recorded %.15f", recorded[p]);
        std::printf("\n");
    }
    if (opt.iterative) {
        std::vector<double> th = paths[first];
        std::vector<double> trace(5 * 10000);
        int ntrace = 0;
        double E = 0;
        if (iterativeOptimise(nparams, th.data(), &E, 0, 10000, trace.data(), &ntrace, ctx) != 0) {
            std::fprintf(stderr, "iterativeOptimise failed\n");
            cleanup(&ctx);
            return 1;
        }
        for (int r = 0; r < ntrace; r++)
            std::printf("iteration %d: %.3f s, energy %.15f, |projected gradient| %.3e\n", r, trace[5 * r],
                        trace[5 * r + 1], trace[5 * r + 2]);
        std::printf("iterative energy %.15f (with nuclear %.15f)\n", E, E + nuc);
        std::printf("angles");
        for (double x : th) std::printf(" %.15f", x);
        std::printf("\n");
    }
    if (opt.subspace && paths.size() > 1) {
        const int k = (int)std::min<size_t>(paths.size() - 1, 9);
        std::vector<double> A;
        for (int i = 1; i <= k; i++) A.insert(A.end(), paths[i].begin(), paths[i].end());
        std::vector<double> ev(k);
        int nullRank = 0;
        if (subspaceDiagonalise(k, nparams, A.data(), ev.data(), nullptr, nullptr, &nullRank, ctx) != 0) {
            std::fprintf(stderr, "subspaceDiagonalise failed\n");
            cleanup(&ctx);
            return 1;
        }
        std::printf("subspace of paths 1..%d: null space rank %d, eigenvalues", k, nullRank);
        for (double e : ev) std::printf(" %.15f", e);
        std::printf("\n");
    }
    if (opt.properties) {
        const size_t p0 = opt.noHFPath ? 1 : 0;
        const int k = (int)(paths.size() - p0);
        const int nov = opt.noLowest ? 0 : opt.overlaps;
        const int cols = 12 + 2 * nov;
        std::vector<double> A, props((size_t)std::max(k, 0) * cols), low(std::max(nov, 1));
        for (size_t i = p0; i < paths.size(); i++) A.insert(A.end(), paths[i].begin(), paths[i].end());
        if (k < 1 || getProperties(k, nparams, A.data(), nov, opt.noHess ? 1 : 0, props.data(), nullptr, nullptr,
                                   low.data(), ctx) != 0) {
            std::fprintf(stderr, "getProperties failed\n");
            cleanup(&ctx);
            return 1;
        }
        if (nov > 0) {
            std::printf("lowest eigenvalues of H:");
            for (int t = 0; t < nov; t++) std::printf(" %.15f", low[t]);
            std::printf("\n");
        }
        static const char *names[12] = {"Elec. Energy", "Elec. + Nuc. Energy", "NegativeHessianEValues",
                                        "NearZeroHessianEValues", "PositiveHessianEValues", "NegativeHessianDiagValues",
                                        "NearZeroHessianDiagValues", "PositiveHessianDiagValues",
                                        "PositiveMetricEValues", "ZeroMetricEValues", "ZeroMetricDiagonalValues",
                                        "NormOfGradVector"};
        std::printf("%-28s", "path");
        for (int i = 0; i < k; i++) std::printf("%22zu", p0 + i);
        std::printf("\n");
        for (int c = 0; c < cols; c++) {
            if (c >= 12 && c < 12 + 2 * nov)
                std::printf("%-28s", (std::string(c % 2 ? "MagOfOverlapWithState_" : "OverlapWithState_") +
                                      std::to_string((c - 12) / 2)).c_str());
            else
                std::printf("%-28s", names[c]);
            if (opt.noHess && c >= 2 && c <= 10) {
                std::printf("%22s\n", "-");
                continue;
            }
            for (int i = 0; i < k; i++) {
                const double v = props[(size_t)i * cols + c];
                if (c >= 2 && c <= 10) std::printf("%22d", (int)v);
                else std::printf("%22.15f", v + (c == 1 ? nuc : 0.0));
            }
            std::printf("\n");
        }
    }
    if (opt.optimise) {
        std::vector<double> th = paths[first];
        std::vector<double> trace(5 * 100000);
        int ntrace = 0;
        double E = 0;
        if (optimise(nparams, th.data(), &E, NAN, 100000, trace.data(), &ntrace, ctx) != 0) {
            std::fprintf(stderr, "optimise failed\n");
            cleanup(&ctx);
            return 1;
        }
        for (int r = 0; r < ntrace; r++)
            std::printf("step %d: %.3f s, energy %.15f, |g| %.3e\n", r, trace[5 * r], trace[5 * r + 1],
                        trace[5 * r + 2]);
        std::printf("optimised energy %.15f (with nuclear %.15f)\n", E, E + nuc);
        std::printf("angles");
        for (double x : th) std::printf(" %.15f", x);
        std::printf("\n");
    }
    cleanup(&ctx);
    return 0;
}
