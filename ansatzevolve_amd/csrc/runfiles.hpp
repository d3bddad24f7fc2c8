// Readers of the reference's run files, named <prefix>_<kind>
// (src/TUPSLoadingUtils.cpp, src/hamiltonianmatrix.cpp:17-135).  Indices
// keep the files' conventions where stated; every reader returns false with
// a message in err when a file is missing or malformed.
#pragma once
#include <array>
#include <cstdint>
#include <string>
#include <utility>
#include <vector>

namespace ae {
namespace files {

// _Initial.dat: lines "bitstring, coefficient[, imaginary]"; the rightmost
// character is bit 0, the string length is the qubit count
// (src/TUPSLoadingUtils.cpp:99-154).  Indices are zero-based full-space.
bool read_initial(const std::string &path, int &numQubits, std::vector<int> &idx, std::vector<double> &coeff,
                  std::string &err);
// _Operators.dat: four one-based spin-orbital indices per line, 0 = absent
// (src/TUPSLoadingUtils.cpp:10-36); returned zero-based, -1 = absent.
bool read_operators(const std::string &path, std::vector<std::array<int8_t, 4>> &ops, std::string &err);
// _Order.dat: per operator "parameter" or "parameter,scale", one-based
// parameter (src/TUPSLoadingUtils.cpp:236-266); returned zero-based.
bool read_order(const std::string &path, std::vector<std::pair<int, double>> &dep, std::string &err);
// _Parameters.dat: sets of "count", "Energy of minimum ... = E ...", then
// the values (src/TUPSLoadingUtils.cpp:226-234).  A set whose value count
// differs from its count line (placeholder sets in some files) is skipped.
bool read_parameter_sets(const std::string &path, std::vector<double> &energies,
                         std::vector<std::vector<double>> &sets, std::string &err);
// _Nuclear_Energy.dat: one number; a missing file gives 0
// (src/TUPSLoadingUtils.cpp:156-174).
bool read_nuclear_energy(const std::string &path, double &e, bool &found, std::string &err);
// The Hamiltonian of a run path, as the reference's file constructor
// (src/hamiltonianmatrix.cpp:660-705) tries them: the text triplets
// <prefix>_Ham_Coeff.dat ("real[, imaginary]" per line) and
// <prefix>_Ham_Index.dat ("i j", one-based full-space) when both exist
// (src/hamiltonianmatrix.cpp:17-79; returned zero-based, kind 1), else the
// integrals <prefix>_oneEInts.bin (n^2 FP64) and _twoEInts.bin (n^4 FP64),
// n = numQubits / 2 (src/hamiltonianmatrix.cpp:402-435; kind 2).
bool read_run_hamiltonian(const std::string &prefix, int numQubits, int &kind, std::vector<int> &i0,
                          std::vector<int> &j0, std::vector<double> &v, std::vector<double> &h,
                          std::vector<double> &g, std::string &err);

} // namespace files
} // namespace ae
