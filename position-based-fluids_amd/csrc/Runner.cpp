// Runner.cpp - headless time loop (src/Runner.cpp:24-160).
#include "Runner.hpp"
#include "WaveGenerator.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <numeric>
#include <sstream>

namespace {

// The .in format of PartReader (count, then "m x y z vx vy vz" lines), with
// 9 significant digits so that reading the file back gives the same floats.
void writeParticles(const std::string &name, unsigned n, const Float4 *pos,
                    const Float4 *vel) {
  std::ofstream o(name.c_str());
  o << std::setprecision(9);
  o << n << "\n";
  for (unsigned i = 0; i < n; ++i)
    o << vel[i].s[3] << " " << pos[i].s[0] << " " << pos[i].s[1] << " "
      << pos[i].s[2] << " " << vel[i].s[0] << " " << vel[i].s[1] << " "
      << vel[i].s[2] << "\n";
}

void writeVtk(const std::string &name, unsigned n, const Float4 *pos,
              const Float4 *vel) {
  std::ofstream o(name.c_str());
  o << std::setprecision(9);
  o << "# vtk DataFile Version 3.0\npbf particles\nASCII\n"
    << "DATASET POLYDATA\nPOINTS " << n << " float\n";
  for (unsigned i = 0; i < n; ++i)
    o << pos[i].s[0] << " " << pos[i].s[1] << " " << pos[i].s[2] << "\n";
  o << "POINT_DATA " << n << "\nVECTORS velocity float\n";
  for (unsigned i = 0; i < n; ++i)
    o << vel[i].s[0] << " " << vel[i].s[1] << " " << vel[i].s[2] << "\n";
}

}  // namespace

std::vector<double> Runner::run(const ConfigParameters &parameters,
                                Simulation &simulation, bool waves,
                                unsigned maxSteps) const {
  const Float4 sizesMin = simulation.getSizesMin();
  const Float4 sizesMax = simulation.getSizesMax();
  simulation.init();
  simulation.initCells();

  float time = 0.0f;
  float wave = 0.0f;
  unsigned stepCount = 0;
  std::vector<double> times;

  do {
    const auto t0 = std::chrono::steady_clock::now();
    simulation.step();
    time += parameters.timeStepLength;
    if (waves) {
      // Runner.cpp:91-104: x-min wall oscillates over a third of the box
      const float push = (sizesMax.s[0] - sizesMin.s[0]) / 3.0f;
      simulation.setWaveGenerator(waveGeneratorValue(push, wave));
      wave += parameters.timeStepLength;
    }
    const auto t1 = std::chrono::steady_clock::now();
    times.push_back(std::chrono::duration<double>(t1 - t0).count());
    ++stepCount;

    const bool partOut =
        parameters.partOutFreq && stepCount % parameters.partOutFreq == 0;
    const bool vtkOut =
        parameters.vtkOutFreq && stepCount % parameters.vtkOutFreq == 0;
    if (partOut || vtkOut) {
      Float4 *pos = nullptr, *vel = nullptr;
      simulation.dumpData(pos, vel);
      const unsigned n = simulation.getNumberParticles();
      if (partOut) {
        std::ostringstream nm;
        nm << parameters.partOutNameBase << stepCount / parameters.partOutFreq
           << ".out";
        writeParticles(nm.str(), n, pos, vel);
      }
      if (vtkOut) {
        std::ostringstream nm;
        nm << parameters.vtkOutNameBase << stepCount / parameters.vtkOutFreq
           << ".vtk";
        writeVtk(nm.str(), n, pos, vel);
      }
    }
    if (maxSteps && stepCount >= maxSteps) break;
  } while (time <= parameters.timeEnd);

  // Runner.cpp:135-150
  std::vector<double> sorted = times;
  const double sum = std::accumulate(times.begin(), times.end(), 0.0);
  const double mean = sum / times.size();
  const double sq = std::inner_product(times.begin(), times.end(),
                                       times.begin(), 0.0);
  const double stdev = std::sqrt(std::max(0.0, sq / times.size() - mean * mean));
  std::sort(sorted.begin(), sorted.end());
  const size_t sz = sorted.size();
  const double median = (sz % 2 == 0)
                            ? (sorted[sz / 2 - 1] + sorted[sz / 2]) / 2
                            : sorted[sz / 2];
  std::cout << "mean: " << mean << std::endl;
  std::cout << "median: " << median << std::endl;
  std::cout << "std: " << stdev << std::endl;
  return times;
}
