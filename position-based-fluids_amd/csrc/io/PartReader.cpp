// PartReader.cpp - reads a .in particle file: a count line, then one
// particle per line "m x y z vx vy vz" (src/io/PartReader.cpp:21-97).
// Error behaviour of the reference: a missing line or an unparsable mass
// stops reading (the particles read so far are returned); an unparsable
// position or velocity component is reported and the particle is kept
// with the components read so far.  The acceleration is zeroed (the
// reference leaves it uninitialised; nothing reads it).
#include "PartReader.hpp"

#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>

std::vector<Particle> PartReader::read(const std::string &filename) const {
  std::vector<Particle> particles;
  std::ifstream ifs(filename.c_str());
  if (!ifs) throw std::runtime_error("Could not open particle file!");

  std::string line;
  std::getline(ifs, line);
  if (!ifs) throw std::runtime_error("Could not parse number of particles!");
  size_t numParticles = 0;
  {
    std::istringstream ssn(line);
    ssn >> numParticles;
  }
  particles.reserve(numParticles);

  for (size_t lineNumber = 0; lineNumber < numParticles; ++lineNumber) {
    if (!ifs) {
      std::cerr << "Error parsing particle in line " << (lineNumber + 2)
                << std::endl;
      break;
    }
    std::getline(ifs, line);
    std::istringstream ss(line);
    Particle q{};
    if (!(ss >> q.m)) {
      std::cerr << "Error parsing particle mass in line " << (lineNumber + 2)
                << std::endl;
      break;
    }
    for (int i = 0; i < DIMENSIONS; ++i) {
      if (!(ss >> q.x[i])) {
        std::cerr << "Error parsing particle position " << i << " in line "
                  << (lineNumber + 2) << std::endl;
        break;
      }
    }
    for (int i = 0; i < DIMENSIONS; ++i) {
      if (!(ss >> q.v[i])) {
        std::cerr << "Error parsing particle velocity " << i << " in line "
                  << (lineNumber + 2) << std::endl;
        break;
      }
    }
    particles.push_back(q);
  }
  return particles;
}
