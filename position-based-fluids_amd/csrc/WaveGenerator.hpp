// WaveGenerator.hpp - the moving x-min wall offset of the driver loop
// (src/Runner.cpp:91-104): a sine over one third of the box in x, one
// period per second of simulated time, starting at offset 0.
#pragma once

#include <cmath>

// Offset for the elapsed wave time `wave` (accumulated in float, as the
// reference does) and push length (x_max - x_min) / 3.  The reference keeps
// frequency and phase as cl_float constants and evaluates
// sin(2.0f * M_PI * f * wave + phase) in double; the phase -pi/2 is
// therefore the float nearest to it, not the exact double.
inline float waveGeneratorValue(float pushLength, float wave) {
  const double pi = 3.14159265358979323846;
  const float frequency = 1.0f;
  const float phase = (float)(-pi / 2.0);
  return (float)(std::sin(2.0f * pi * frequency * wave + phase) * pushLength /
                     2.0f +
                 pushLength / 2.0f);
}
