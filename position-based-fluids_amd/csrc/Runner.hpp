// Runner.hpp - headless driver loop of the reference (src/Runner.hpp:11-23).
#pragma once

#include <vector>

#include "Parameters.hpp"
#include "Simulation.hpp"

class Runner {
 private:
  Runner &operator=(const Runner &other);
  Runner(const Runner &other);

 public:
  Runner() {}

  // do { step; time += dt; wave } while (time <= timeEnd)
  // (src/Runner.cpp:81-133), then mean / median / std of the step times
  // (:135-150) on stdout.  Rendering is replaced by optional particle
  // (.out, PartReader format) and legacy VTK output every partOutFreq /
  // vtkOutFreq steps.  waves: drive the moving wall as the 'G' key of the
  // reference does (:91-104).  Returns the per-step wall times (seconds).
  std::vector<double> run(const ConfigParameters &parameters,
                          Simulation &simulation, bool waves = false,
                          unsigned maxSteps = 0) const;
};
