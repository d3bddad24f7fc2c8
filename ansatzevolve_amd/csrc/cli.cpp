// Command-line front end over the C ABI, with the argument words of the
// reference's cppAnsatzSynth (src/main.cpp:93-219):
//
//   ansatzsynth filepath <prefix> [optimise] [benchmark] [onlyevolve]
//               [loadpaths N] [help]
//
// It reads the run files <prefix>_* (start state, operators, order file,
// Hamiltonian as text triplets or integrals, nuclear energy, parameter
// sets) and, like the reference, takes path 0 as all angles zero and path
// i >= 1 as the i-th parameter set (src/TUPSLoadingUtils.cpp:226-266).
// Without a mode it prints the energy of every path; optimise runs the
// Newton optimiser from path 1 (src/main.cpp:488-499); benchmark times
// energy + gradient evaluations at path 1 (src/main.cpp:396-425);
// onlyevolve evolves path 1 and prints the final state's norm
// (src/main.cpp:380-392).  Energies are printed electronic and with the
// nuclear energy.
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
        "  loadpaths N        keep the first N paths (path 0 included)\n");
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
