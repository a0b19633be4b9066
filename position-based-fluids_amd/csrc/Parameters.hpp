// Parameters.hpp - ConfigParameters of the reference (src/Parameters.hpp:12-31)
// plus the physics settings the reference hard-codes in its kernels.
#pragma once

#include <string>

#include "../../include/pbf.h"

struct ConfigParameters {
  // src/Parameters.hpp:13-30, same names and meaning.  The reference leaves
  // keys missing from a .par file uninitialised; here they default to 0
  // (and the physics block to the REF defaults of pbf_params_init).
  std::string partInputFile;
  float timeStepLength = 0.0f;
  float timeEnd = 0.0f;
  unsigned partOutFreq = 0;
  std::string partOutNameBase;
  unsigned vtkOutFreq = 0;
  std::string vtkOutNameBase;
  unsigned clWorkGroupSize1D = 0;
  float xMin = 0.0f, xMax = 0.0f;
  float yMin = 0.0f, yMax = 0.0f;
  float zMin = 0.0f, zMax = 0.0f;
  float xN = 0.0f, yN = 0.0f, zN = 0.0f;
  float restDensity = 0.0f;

  // Physics block (physics mode, iterations, relaxation, s_corr, XSPH,
  // vorticity, gravity, precision, obstacle, block size); see pbf.h.
  pbf_params physics;

  ConfigParameters() { pbf_params_init(&physics, PBF_PHYSICS_REF); }

  // The C-ABI parameter block of this configuration.
  pbf_params toPbfParams() const {
    pbf_params p = physics;
    p.time_step_length = timeStepLength;
    p.time_end = timeEnd;
    p.x_min = xMin;
    p.x_max = xMax;
    p.y_min = yMin;
    p.y_max = yMax;
    p.z_min = zMin;
    p.z_max = zMax;
    p.x_n = xN;
    p.y_n = yN;
    p.z_n = zN;
    p.rest_density = restDensity;
    return p;
  }
};
