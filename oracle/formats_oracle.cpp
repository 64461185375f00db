// formats_oracle.cpp -- CPU restatement of GLAMER's csv particle reader.
//
// TEST INFRASTRUCTURE ONLY (the checker and the timed CPU baseline of the
// csv ingest); the product library never links or calls it.
//
// Restates MakeParticleLenses::readCSV (MultiPlane/particle_lens.cpp:211-342)
// for 3, 4 and 5 columns the way the reference runs it: one thread, two passes
// over an std::ifstream with std::getline, leading lines that start with '#'
// skipped, every line split at ',' as Utilities::splitstring does
// (TreeCode_link/utilities.cpp:828-837: the last field is the rest of the
// line) and converted with std::stof, whose exceptions end the read.

#include <cstddef>
#include <cstdint>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

void split_commas(std::string line, std::vector<std::string>& vec) {
  size_t pos;
  while ((pos = line.find(',')) != std::string::npos) {
    vec.push_back(line.substr(0, pos));
    line.erase(0, pos + 1);
  }
  vec.push_back(line);
}

}  // namespace

extern "C" {

// Returns 0 and fills out[5*i + (x, y, z, Mass, Size)] for i < *n (when
// out != nullptr and cap >= *n); 1: the first data line has too few columns;
// 2: stof found nothing to convert; 3: stof out of range; 4: a later line has
// fewer fields than needed (undefined in the reference).  *bad_line is the
// 0-based data line of a 2/3/4.
int orc_csv_read(const char* path, int columns, float* out, size_t cap, size_t* n,
                 size_t* bad_line) {
  std::ifstream file(path);
  std::string line;
  while (std::getline(file, line) && line[0] == '#') {
  }
  std::vector<std::string> vec;
  split_commas(line, vec);
  if ((int)vec.size() < columns) return 1;
  size_t ntot = 1;
  while (std::getline(file, line)) ++ntot;
  *n = ntot;
  if (!out || cap < ntot) return 0;
  file.clear();
  file.seekg(0);
  while (std::getline(file, line) && line[0] == '#') {
  }
  size_t k = 0;
  do {
    std::vector<std::string> f;
    split_commas(line, f);
    if ((int)f.size() < columns) {
      *bad_line = k;
      return 4;
    }
    float v[5] = {0, 0, 0, 1.0f, 0};
    try {
      for (int c = 0; c < columns; ++c) v[c] = std::stof(f[c]);
    } catch (const std::invalid_argument&) {
      *bad_line = k;
      return 2;
    } catch (const std::out_of_range&) {
      *bad_line = k;
      return 3;
    }
    for (int c = 0; c < 5; ++c) out[5 * k + c] = v[c];
    ++k;
  } while (std::getline(file, line));
  return 0;
}

}  // extern "C"
