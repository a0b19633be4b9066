// ConfigReader.hpp - .par reader (src/io/ConfigReader.hpp:11-24).
#pragma once

#include <string>

#include "../Parameters.hpp"

class ConfigReader {
 private:
  ConfigReader &operator=(const ConfigReader &other);
  ConfigReader(const ConfigReader &other);

 public:
  ConfigReader() {}

  ConfigParameters read(const std::string &filename) const;
};
