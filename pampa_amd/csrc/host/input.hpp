// Reader of the `.pmp` keyword format (line cleaning, tokenising and typed
// readers), following input.cxx:4-427 of the reference.
#pragma once

#include <climits>
#include <cfloat>

#include "common.hpp"

namespace pampa {
namespace input {

// Trim, turn a tab and the character after it into one space, squeeze double
// spaces and drop '#' comment lines (input.cxx:4-38).
void clean(std::string& s);

// Next non-empty line split on single spaces (input.cxx:41-60).
std::vector<std::string> next_line(std::istream& in);

int read(int& x, int lo, int hi, const std::string& s);
int read(double& x, double lo, double hi, const std::string& s);
int read(bool& b, const std::string& s);
// n values spread over as many lines as needed
int read(std::vector<double>& v, int n, double lo, double hi, std::istream& in);
int read(std::vector<int>& v, int n, int lo, int hi, std::istream& in);
// n lines of exactly m values
int read_rows(std::vector<double>& v, int n, int m, double lo, double hi, std::istream& in);
// n lines of any length, total at most nt (ragged rows)
int read_ragged(std::vector<int>& ptr, std::vector<int>& v, int n, int nt, int lo, int hi,
                std::istream& in);
// function: single value, or "<n> {" with time and value tables (input.cxx:268-311)
int read(Function& f, double lo, double hi, const std::vector<std::string>& line, unsigned& i,
         std::istream& in);
// boundary condition (input.cxx:367-404)
int read(BoundaryCondition& bc, const std::vector<std::string>& line, unsigned& i, std::istream& in);

// Open a file named in an input: relative to the working directory first (the
// reference's behaviour), then relative to the directory of the including file.
bool open(std::ifstream& f, const std::string& name, const std::string& base_dir);
std::string dir_of(const std::string& path);

}  // namespace input
}  // namespace pampa
