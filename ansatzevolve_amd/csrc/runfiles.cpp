// Readers of the reference's run files (see runfiles.hpp).
#include "runfiles.hpp"

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

namespace ae {
namespace files {

namespace {

bool read_lines(const std::string &path, std::vector<std::string> &lines, std::string &err)
{
    std::ifstream f(path);
    if (!f) {
        err = "cannot open " + path;
        return false;
    }
    lines.clear();
    std::string s;
    while (std::getline(f, s)) {
        // trim (files come with trailing spaces and Windows line ends)
        size_t a = 0, b = s.size();
        while (a < b && std::isspace((unsigned char)s[a])) a++;
        while (b > a && std::isspace((unsigned char)s[b - 1])) b--;
        if (b > a) lines.push_back(s.substr(a, b - a));
    }
    return true;
}

// every number of a line, separators being commas and white space
bool numbers(const std::string &line, std::vector<double> &out)
{
    out.clear();
    const char *p = line.c_str();
    while (*p) {
        while (*p && (std::isspace((unsigned char)*p) || *p == ',')) p++;
        if (!*p) break;
        char *e = nullptr;
        const double v = std::strtod(p, &e);
        if (e == p) return false;
        out.push_back(v);
        p = e;
    }
    return true;
}

bool file_exists(const std::string &path)
{
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    std::fclose(f);
    return true;
}

bool read_binary(const std::string &path, size_t count, std::vector<double> &out, std::string &err)
{
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) {
        err = "cannot open " + path;
        return false;
    }
    std::fseek(f, 0, SEEK_END);
    const long sz = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    if (sz < 0 || (size_t)sz != count * sizeof(double)) {
        std::fclose(f);
        err = path + " holds " + std::to_string(sz) + " bytes, expected " + std::to_string(count * sizeof(double)) +
              " (molecular-orbital integrals without permutational packing)";
        return false;
    }
    out.resize(count);
    const size_t got = std::fread(out.data(), sizeof(double), count, f);
    std::fclose(f);
    if (got != count) {
        err = "short read of " + path;
        return false;
    }
    return true;
}

} // namespace

bool read_initial(const std::string &path, int &numQubits, std::vector<int> &idx, std::vector<double> &coeff,
                  std::string &err)
{
    std::vector<std::string> lines;
    if (!read_lines(path, lines, err)) return false;
    numQubits = 0;
    idx.clear();
    coeff.clear();
    for (const std::string &l : lines) {
        size_t k = 0;
        while (k < l.size() && (l[k] == '0' || l[k] == '1')) k++;
        if (k == 0 || k > 31) {
            err = path + ": expected a bitstring of 1 to 31 characters: " + l;
            return false;
        }
        int bits = 0;
        for (size_t i = 0; i < k; i++)
            if (l[k - 1 - i] == '1') bits |= 1 << i;
        std::vector<double> v;
        if (!numbers(l.substr(k), v)) {
            err = path + ": unreadable coefficient: " + l;
            return false;
        }
        // a bitstring alone means coefficient 1 (src/TUPSLoadingUtils.cpp:121-126)
        const double c = v.empty() ? 1.0 : v[0];
        if (v.size() > 1 && v[1] != 0.0) std::fprintf(stderr, "Imaginary coeff not zero: %s\n", l.c_str());
        numQubits = std::max(numQubits, (int)k);
        idx.push_back(bits);
        coeff.push_back(c);
    }
    if (idx.empty()) {
        err = path + ": no basis state";
        return false;
    }
    return true;
}

bool read_operators(const std::string &path, std::vector<std::array<int8_t, 4>> &ops, std::string &err)
{
    std::vector<std::string> lines;
    if (!read_lines(path, lines, err)) return false;
    ops.clear();
    for (const std::string &l : lines) {
        std::vector<double> v;
        if (!numbers(l, v) || v.size() != 4) {
            err = path + ": expected four indices: " + l;
            return false;
        }
        std::array<int8_t, 4> o;
        for (int j = 0; j < 4; j++) {
            const int x = (int)v[j];
            if (x < 0 || x > 128 || (double)x != v[j]) {
                err = path + ": index out of range: " + l;
                return false;
            }
            o[j] = (int8_t)(x - 1);
        }
        ops.push_back(o);
    }
    if (ops.empty()) {
        err = path + ": no operators";
        return false;
    }
    return true;
}

bool read_order(const std::string &path, std::vector<std::pair<int, double>> &dep, std::string &err)
{
    std::vector<std::string> lines;
    if (!read_lines(path, lines, err)) return false;
    dep.clear();
    for (const std::string &l : lines) {
        std::vector<double> v;
        if (!numbers(l, v) || v.empty() || v.size() > 2 || v[0] < 1 || (double)(int)v[0] != v[0]) {
            err = path + ": expected \"parameter\" or \"parameter,scale\" (one-based): " + l;
            return false;
        }
        dep.push_back({(int)v[0] - 1, v.size() == 2 ? v[1] : 1.0});
    }
    if (dep.empty()) {
        err = path + ": no entries";
        return false;
    }
    return true;
}

bool read_parameter_sets(const std::string &path, std::vector<double> &energies,
                         std::vector<std::vector<double>> &sets, std::string &err)
{
    std::vector<std::string> lines;
    if (!read_lines(path, lines, err)) return false;
    energies.clear();
    sets.clear();
    auto is_count = [](const std::string &s) {
        if (s.empty()) return false;
        for (char c : s)
            if (!std::isdigit((unsigned char)c)) return false;
        return true;
    };
    std::vector<size_t> starts;
    for (size_t i = 0; i + 1 < lines.size(); i++)
        if (is_count(lines[i]) && lines[i + 1].compare(0, 6, "Energy") == 0) starts.push_back(i);
    for (size_t s = 0; s < starts.size(); s++) {
        const size_t a = starts[s], b = s + 1 < starts.size() ? starts[s + 1] : lines.size();
        const long count = std::atol(lines[a].c_str());
        const size_t eq = lines[a + 1].find('=');
        if (eq == std::string::npos) {
            err = path + ": energy line without '=': " + lines[a + 1];
            return false;
        }
        char *e = nullptr;
        const char *p = lines[a + 1].c_str() + eq + 1;
        const double E = std::strtod(p, &e);
        if (e == p) {
            err = path + ": unreadable energy: " + lines[a + 1];
            return false;
        }
        std::vector<double> vals;
        for (size_t i = a + 2; i < b; i++) {
            std::vector<double> v;
            if (!numbers(lines[i], v)) {
                err = path + ": unreadable value: " + lines[i];
                return false;
            }
            vals.insert(vals.end(), v.begin(), v.end());
        }
        if ((long)vals.size() != count) continue;  // placeholder set
        energies.push_back(E);
        sets.push_back(std::move(vals));
    }
    return true;
}

bool read_nuclear_energy(const std::string &path, double &e, bool &found, std::string &err)
{
    e = 0;
    found = file_exists(path);
    if (!found) return true;
    std::vector<std::string> lines;
    if (!read_lines(path, lines, err)) return false;
    std::vector<double> v;
    if (lines.empty() || !numbers(lines[0], v) || v.empty()) {
        err = path + ": expected one number";
        return false;
    }
    e = v[0];
    return true;
}

bool read_run_hamiltonian(const std::string &prefix, int numQubits, int &kind, std::vector<int> &i0,
                          std::vector<int> &j0, std::vector<double> &v, std::vector<double> &h,
                          std::vector<double> &g, std::string &err)
{
    kind = 0;
    const std::string coeffPath = prefix + "_Ham_Coeff.dat", indexPath = prefix + "_Ham_Index.dat";
    if (file_exists(coeffPath) && file_exists(indexPath)) {
        std::vector<std::string> cl, il;
        if (!read_lines(coeffPath, cl, err) || !read_lines(indexPath, il, err)) return false;
        // pairs are read until either file ends (src/hamiltonianmatrix.cpp:45-73)
        const size_t n = std::min(cl.size(), il.size());
        i0.clear();
        j0.clear();
        v.clear();
        for (size_t k = 0; k < n; k++) {
            std::vector<double> c, ij;
            if (!numbers(cl[k], c) || c.empty() || !numbers(il[k], ij) || ij.size() != 2 || ij[0] < 1 || ij[1] < 1) {
                err = prefix + ": unreadable Hamiltonian entry " + std::to_string(k + 1);
                return false;
            }
            i0.push_back((int)ij[0] - 1);
            j0.push_back((int)ij[1] - 1);
            v.push_back(c[0]);
        }
        if (v.empty()) {
            err = prefix + ": empty text Hamiltonian";
            return false;
        }
        kind = 1;
        return true;
    }
    if (numQubits < 2 || numQubits % 2) {
        err = prefix + ": integrals need an even qubit count";
        return false;
    }
    const size_t n = (size_t)numQubits / 2;
    std::string e1, e2;
    if (!read_binary(prefix + "_oneEInts.bin", n * n, h, e1)) {
        err = "no text Hamiltonian (" + coeffPath + ", " + indexPath + ") and " + e1;
        return false;
    }
    if (!read_binary(prefix + "_twoEInts.bin", n * n * n * n, g, e2)) {
        err = e2;
        return false;
    }
    kind = 2;
    return true;
}

} // namespace files
} // namespace ae
