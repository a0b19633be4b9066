// pbf_run.cpp - headless counterpart of the reference's main()
// (src/main.cpp:40-189): read a .par scenario and its particle file, build
// the Simulation on the current HIP device and run the Runner loop.
//   pbf_run <scenario.par> [--steps N] [--waves]
#include <cstdlib>
#include <cstring>
#include <exception>
#include <iostream>
#include <string>

#include "Runner.hpp"
#include "Simulation.hpp"
#include "io/ConfigReader.hpp"
#include "io/PartReader.hpp"

int main(int argc, char **argv) {
  if (argc < 2) {
    std::cerr << "usage: pbf_run <scenario.par> [--steps N] [--waves]\n";
    return 2;
  }
  unsigned steps = 0;
  bool waves = false;
  for (int a = 2; a < argc; ++a) {
    if (!std::strcmp(argv[a], "--steps") && a + 1 < argc)
      steps = (unsigned)std::strtoul(argv[++a], nullptr, 10);
    else if (!std::strcmp(argv[a], "--waves"))
      waves = true;
  }
  try {
    const std::string par = argv[1];
    ConfigReader configReader;
    ConfigParameters parameters = configReader.read(par);
    std::string dir;
    const size_t slash = par.find_last_of('/');
    if (slash != std::string::npos) dir = par.substr(0, slash + 1);
    PartReader partReader;
    const std::vector<Particle> particles =
        partReader.read(dir + parameters.partInputFile);
    Simulation simulation(parameters, particles);
    Runner runner;
    runner.run(parameters, simulation, waves, steps);
  } catch (const std::exception &e) {
    std::cerr << "STD Error caught: " << e.what() << std::endl;
    return 1;
  }
  return 0;
}
