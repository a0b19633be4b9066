// ConfigReader.cpp - reads a .par scenario file: one "key value" pair per
// line, the keys of src/io/ConfigReader.cpp:45-84, unknown keys reported on
// stderr and skipped (:81-83).  The physics keys below have no reference
// counterpart (the reference hard-codes those values in its kernels).
#include "ConfigReader.hpp"

#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <vector>

ConfigParameters ConfigReader::read(const std::string &filename) const {
  ConfigParameters p;
  std::ifstream ifs(filename.c_str());
  if (!ifs) throw std::runtime_error("Could not open parameter file!");

  std::vector<std::string> lines;
  std::string line, key;
  while (std::getline(ifs, line)) lines.push_back(line);

  // The physics mode sets the defaults of the whole physics block, so it is
  // applied before every other key wherever it stands in the file (the last
  // one wins); the explicit physics keys then override those defaults.
  for (const std::string &l : lines) {
    std::istringstream ss(l);
    std::string v;
    if (!(ss >> key) || key != "physics") continue;
    ss >> v;
    if (v == "ref")
      pbf_params_init(&p.physics, PBF_PHYSICS_REF);
    else if (v == "paper")
      pbf_params_init(&p.physics, PBF_PHYSICS_PAPER);
    else
      throw std::runtime_error("physics must be 'ref' or 'paper'");
  }

  for (const std::string &l : lines) {
    line = l;
    if (!line.empty()) {
      std::istringstream ss(line);
      if (!(ss >> key)) {
        std::cerr << "Unable to read parameter" << std::endl;
      } else if (key == "part_input_file") {
        ss >> p.partInputFile;
      } else if (key == "timestep_length") {
        ss >> p.timeStepLength;
      } else if (key == "time_end") {
        ss >> p.timeEnd;
      } else if (key == "part_out_freq") {
        ss >> p.partOutFreq;
      } else if (key == "part_out_name_base") {
        ss >> p.partOutNameBase;
      } else if (key == "vtk_out_freq") {
        ss >> p.vtkOutFreq;
      } else if (key == "vtk_out_name_base") {
        ss >> p.vtkOutNameBase;
      } else if (key == "cl_workgroup_1dsize") {
        ss >> p.clWorkGroupSize1D;
      } else if (key == "x_min") {
        ss >> p.xMin;
      } else if (key == "x_max") {
        ss >> p.xMax;
      } else if (key == "y_min") {
        ss >> p.yMin;
      } else if (key == "y_max") {
        ss >> p.yMax;
      } else if (key == "z_min") {
        ss >> p.zMin;
      } else if (key == "z_max") {
        ss >> p.zMax;
      } else if (key == "x_n") {
        ss >> p.xN;
      } else if (key == "y_n") {
        ss >> p.yN;
      } else if (key == "z_n") {
        ss >> p.zN;
      } else if (key == "restdensity") {
        ss >> p.restDensity;
      } else if (key == "physics") {
        // applied above, before the other keys
      } else if (key == "iterations") {
        ss >> p.physics.iterations;
      } else if (key == "relaxation") {
        ss >> p.physics.relaxation;
      } else if (key == "scorr_k") {
        ss >> p.physics.scorr_k;
      } else if (key == "scorr_n") {
        ss >> p.physics.scorr_n;
      } else if (key == "scorr_dq") {
        ss >> p.physics.scorr_dq;
      } else if (key == "xsph_c") {
        ss >> p.physics.xsph_c;
      } else if (key == "vorticity_eps") {
        ss >> p.physics.vorticity_eps;
      } else if (key == "gravity_y") {
        ss >> p.physics.gravity_y;
      } else if (key == "precision") {
        std::string v;
        ss >> v;
        if (v == "fp32")
          p.physics.precision = PBF_PRECISION_FP32;
        else if (v == "mixed")
          p.physics.precision = PBF_PRECISION_MIXED;
        else
          throw std::runtime_error("precision must be 'fp32' or 'mixed'");
      } else if (key == "fp32_iterations") {
        ss >> p.physics.fp32_iterations;
      } else if (key == "sphere") {
        for (int a = 0; a < 4; ++a) ss >> p.physics.sphere[a];
      } else if (key == "block_size") {
        ss >> p.physics.block_size;
      } else if (key == "launch_grid") {
        ss >> p.physics.launch_grid;
      } else if (key == "scan_cells") {
        ss >> p.physics.scan_cells;
      } else {
        std::cerr << "Unknown parameter " << key << std::endl
                  << "Leaving it out." << std::endl;
      }
    }
  }
  return p;
}
