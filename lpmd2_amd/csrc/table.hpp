// Pair table of the tabulatedpair plugin (plugins/tabulatedpair.cc:34-79),
// host-only: rows of exactly three whitespace-separated fields (r, E, ff) whose
// first non-blank character is not '#', in file order; the cutoff is the last
// r larger than every r before it (the reference's "previo" rule), bins the
// number of rows.  Fields are read with atof, as the reference does.
#pragma once

#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

namespace lpmdx {

struct PairTable {
  std::vector<double> e, f;
  double cutoff = 0.0;
  bool ok = false; // the file opened
};

inline PairTable read_pair_table(const std::string& path) {
  PairTable T;
  std::ifstream in(path);
  if (!in.is_open())
    return T;
  T.ok = true;
  double previo = 0.0;
  std::string line;
  while (std::getline(in, line)) {
    // util.h IsSpace: ' ', '\n', '\t', '\r' separate fields
    for (char& c : line)
      if (c == '\t' || c == '\r' || c == '\n')
        c = ' ';
    std::istringstream ws(line);
    std::vector<std::string> w;
    std::string tok;
    while (ws >> tok)
      w.push_back(tok);
    if (w.size() != 3 || w[0][0] == '#')
      continue;
    const double r = std::atof(w[0].c_str());
    if (r > previo) {
      T.cutoff = r;
      previo = r;
    }
    T.e.push_back(std::atof(w[1].c_str()));
    T.f.push_back(std::atof(w[2].c_str()));
  }
  return T;
}

} // namespace lpmdx
