// PartReader.hpp - .in particle reader (src/io/PartReader.hpp:13-26).
#pragma once

#include <string>
#include <vector>

#include "../Particle.hpp"

class PartReader {
 private:
  PartReader &operator=(PartReader &other);
  PartReader(const PartReader &other);

 public:
  PartReader() {}

  std::vector<Particle> read(const std::string &filename) const;
};
