// Particle.hpp - input particle record of the reference (src/Particle.hpp:9-14).
#pragma once

static const int DIMENSIONS = 3;

struct Particle {
  float m;              // mass
  float x[DIMENSIONS];  // position
  float v[DIMENSIONS];  // velocity
  float a[DIMENSIONS];  // acceleration (unused by the solver)
};

static_assert(sizeof(Particle) == 40, "Particle must match pbf_particle");

// cl_float4 stand-in for dumpData / getSizesMin (component access via .s[]).
struct Float4 {
  float s[4];
};
