#include "input.hpp"

#include <cstdlib>
#include <cstring>

namespace pampa {

void report_error(const char* file, int line, const char* func, const std::string& msg) {
   const char* base = std::strrchr(file, '/');
   base = base ? base + 1 : file;
   std::string m = msg;
   while (!m.empty() && m.back() == '.') m.pop_back();
   std::cout << base << ":" << line << ":" << func << "(): error: " << m << "." << std::endl;
}

double Function::operator()(double x) const {
   if (t.empty() || x < t.front()) return y.front();
   if (x > t.back()) return y.back();
   size_t i2 = 0;
   while (i2 < t.size() && t[i2] < x) i2++;
   if (i2 == 0) return y.front();
   size_t i1 = i2 - 1;
   double f = (x - t[i1]) / (t[i2] - t[i1]);
   return (1.0 - f) * y[i1] + f * y[i2];
}

namespace input {

void clean(std::string& s) {
   if (s.empty()) return;
   size_t i;
   while ((i = s.find('\t')) != std::string::npos) s.replace(i, 2, " ");
   while ((i = s.find("  ")) != std::string::npos) s.erase(i, 1);
   if (s == " ") {
      s.clear();
      return;
   }
   size_t a = s.find_first_not_of(' '), b = s.find_last_not_of(' ');
   if (a == std::string::npos) {
      s.clear();
      return;
   }
   s = s.substr(a, b - a + 1);
   if (!s.empty() && s[0] == '#') s.clear();
}

std::vector<std::string> next_line(std::istream& in) {
   std::string line;
   std::vector<std::string> words;
   while (std::getline(in, line)) {
      clean(line);
      if (line.empty()) continue;
      size_t p = 0;
      while (true) {
         size_t q = line.find(' ', p);
         words.push_back(line.substr(p, q == std::string::npos ? std::string::npos : q - p));
         if (q == std::string::npos) break;
         p = q + 1;
      }
      break;
   }
   return words;
}

// Leading-number parse, as std::stoi / std::stod do.
int read(int& x, int lo, int hi, const std::string& s) {
   char* end = nullptr;
   long v = std::strtol(s.c_str(), &end, 10);
   PAMPA_CHECK(end == s.c_str(), "wrong int value '" + s + "'");
   x = (int)v;
   PAMPA_CHECK(x < lo || x > hi, "out-of-bounds int value");
   return 0;
}

int read(double& x, double lo, double hi, const std::string& s) {
   char* end = nullptr;
   x = std::strtod(s.c_str(), &end);
   PAMPA_CHECK(end == s.c_str(), "wrong double value '" + s + "'");
   PAMPA_CHECK(x < lo || x > hi, "out-of-bounds double value");
   return 0;
}

int read(bool& b, const std::string& s) {
   int x = 0;
   PAMPA_CHECK(read(x, INT_MIN, INT_MAX, s), "wrong bool value");
   PAMPA_CHECK(x != 0 && x != 1, "wrong bool value");
   b = x;
   return 0;
}

int read(std::vector<double>& v, int n, double lo, double hi, std::istream& in) {
   v.assign(n, 0.0);
   int l = 0;
   while (l < n) {
      auto line = next_line(in);
      PAMPA_CHECK(line.empty(), "missing data");
      for (auto& w : line) {
         PAMPA_CHECK(l >= n, "out-of-bounds data");
         PAMPA_CHECK(read(v[l++], lo, hi, w), "wrong data");
      }
   }
   return 0;
}

int read(std::vector<int>& v, int n, int lo, int hi, std::istream& in) {
   v.assign(n, 0);
   int l = 0;
   while (l < n) {
      auto line = next_line(in);
      PAMPA_CHECK(line.empty(), "missing data");
      for (auto& w : line) {
         PAMPA_CHECK(l >= n, "out-of-bounds data");
         PAMPA_CHECK(read(v[l++], lo, hi, w), "wrong data");
      }
   }
   return 0;
}

int read_rows(std::vector<double>& v, int n, int m, double lo, double hi, std::istream& in) {
   v.assign((size_t)n * m, 0.0);
   for (int l = 0; l < n; l++) {
      auto line = next_line(in);
      PAMPA_CHECK(line.empty(), "missing data");
      PAMPA_CHECK((int)line.size() < m, "missing data");
      PAMPA_CHECK((int)line.size() > m, "out-of-bounds data");
      for (int i = 0; i < m; i++) PAMPA_CHECK(read(v[(size_t)l * m + i], lo, hi, line[i]), "wrong data");
   }
   return 0;
}

int read_ragged(std::vector<int>& ptr, std::vector<int>& v, int n, int nt, int lo, int hi,
                std::istream& in) {
   ptr.assign(1, 0);
   v.clear();
   for (int l = 0; l < n; l++) {
      auto line = next_line(in);
      PAMPA_CHECK(line.empty(), "missing data");
      PAMPA_CHECK((long)v.size() + (long)line.size() > nt, "out-of-bounds data");
      for (auto& w : line) {
         int x = 0;
         PAMPA_CHECK(read(x, lo, hi, w), "wrong data");
         v.push_back(x);
      }
      ptr.push_back((int)v.size());
   }
   return 0;
}

int read(Function& f, double lo, double hi, const std::vector<std::string>& line, unsigned& i,
         std::istream& in) {
   if (line.size() == i + 1) {
      double x = 0.0;
      PAMPA_CHECK(read(x, lo, hi, line[i]), "wrong function data");
      f.t.clear();
      f.y.assign(1, x);
   } else if (line.size() == i + 2) {
      int n = 0;
      PAMPA_CHECK(read(n, 1, INT_MAX, line[i]), "wrong number of time points");
      PAMPA_CHECK(line[++i] != "{", "missing opening '{' for function data");
      PAMPA_CHECK(read(f.t, n, 0.0, DBL_MAX, in), "wrong time data");
      PAMPA_CHECK(read(f.y, n, lo, hi, in), "wrong function data");
      auto close = next_line(in);
      PAMPA_CHECK(close.empty() || close[0] != "}", "missing closing '}' for function data");
   } else {
      PAMPA_CHECK(true, "wrong number of function arguments");
   }
   return 0;
}

int read(BoundaryCondition& bc, const std::vector<std::string>& line, unsigned& i, std::istream& in) {
   PAMPA_CHECK(i >= line.size(), "wrong boundary-condition type");
   const std::string& t = line[i++];
   if (t == "vacuum")
      bc.type = BC_VACUUM;
   else if (t == "reflective")
      bc.type = BC_REFLECTIVE;
   else if (t == "robin")
      bc.type = BC_ROBIN;
   else if (t == "dirichlet")
      bc.type = BC_DIRICHLET;
   else if (t == "adiabatic")
      bc.type = BC_ADIABATIC;
   else if (t == "convection")
      bc.type = BC_CONVECTION;
   else
      PAMPA_CHECK(true, "wrong boundary-condition type");
   if (bc.type == BC_ROBIN || bc.type == BC_DIRICHLET) {
      bc.f.resize(1);
      PAMPA_CHECK(read(bc.f[0], -DBL_MAX, DBL_MAX, line, i, in), "wrong boundary-condition parameter");
   }
   if (bc.type == BC_CONVECTION) {
      // two single values, or a time table with two value rows (input.cxx:314-364)
      bc.f.resize(2);
      if (line.size() == i + 2 && line[i + 1] != "{") {
         for (int j = 0; j < 2; j++) {
            double x = 0.0;
            PAMPA_CHECK(read(x, -DBL_MAX, DBL_MAX, line[i++]), "wrong boundary-condition parameters");
            bc.f[j].y.assign(1, x);
         }
      } else {
         int n = 0;
         PAMPA_CHECK(line.size() != i + 2, "wrong number of function arguments");
         PAMPA_CHECK(read(n, 1, INT_MAX, line[i]), "wrong number of time points");
         PAMPA_CHECK(line[++i] != "{", "missing opening '{' for function data");
         std::vector<double> t;
         PAMPA_CHECK(read(t, n, 0.0, DBL_MAX, in), "wrong time data");
         for (int j = 0; j < 2; j++) {
            bc.f[j].t = t;
            PAMPA_CHECK(read(bc.f[j].y, n, -DBL_MAX, DBL_MAX, in), "wrong function data");
         }
         auto close = next_line(in);
         PAMPA_CHECK(close.empty() || close[0] != "}", "missing closing '}' for function data");
      }
   }
   return 0;
}

std::string dir_of(const std::string& path) {
   size_t p = path.find_last_of('/');
   return p == std::string::npos ? std::string(".") : path.substr(0, p);
}

bool open(std::ifstream& f, const std::string& name, const std::string& base_dir) {
   f.open(name);
   if (f.is_open()) return true;
   if (!name.empty() && name[0] != '/' && !base_dir.empty()) {
      f.clear();
      f.open(base_dir + "/" + name);
   }
   return f.is_open();
}

}  // namespace input
}  // namespace pampa
