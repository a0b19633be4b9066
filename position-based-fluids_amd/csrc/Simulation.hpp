// Simulation.hpp - the reference's solver class (src/Simulation.hpp:30-197)
// over the C ABI of pbf.h.  The OpenCL context/device/kernel map and the GL
// sharing buffer of the reference constructor are replaced by the current
// HIP device; init() and initCells() keep the reference call sequence.
#pragma once

#include <vector>

#include "../../include/pbf.h"
#include "Parameters.hpp"
#include "Particle.hpp"

class Simulation {
 private:
  Simulation &operator=(const Simulation &other);
  Simulation(const Simulation &other);

 public:
  explicit Simulation(const ConfigParameters &parameters,
                      const std::vector<Particle> &particles);
  ~Simulation();

  void init(void);       // allocates and uploads (Simulation.cpp:104-201)
  void initCells(void);  // no-op: cell tables are rebuilt every step
  void step(void);       // one step; returns when it is done (:455-576)

  // Copy current positions and velocities (original particle order).  The
  // pointers stay valid until the next dumpData or destruction (:578-591).
  void dumpData(Float4 *(&positions), Float4 *(&velocities));

  unsigned getNumberParticles() const { return mNumParticles; }
  float getSizeXmin() const { return mSystemSizeMin.s[0]; }
  float getSizeXmax() const { return mSystemSizeMax.s[0]; }
  float getSizeYmin() const { return mSystemSizeMin.s[1]; }
  float getSizeYmax() const { return mSystemSizeMax.s[1]; }
  float getSizeZmin() const { return mSystemSizeMin.s[2]; }
  float getSizeZmax() const { return mSystemSizeMax.s[2]; }
  const Float4 getSizesMin() const { return mSystemSizeMin; }
  const Float4 getSizesMax() const { return mSystemSizeMax; }

  void setWaveGenerator(const double value);

  // The positions buffer the reference shares with its GL viewer (the
  // GLuint of its constructor, src/Simulation.cpp:149-158): device memory of
  // getNumberParticles() float4, e.g. a GL buffer registered with
  // hipGraphicsGLRegisterBuffer and mapped by the caller.  Once set, every
  // step() leaves each particle's (x, |v|) at its index there, as
  // updatePositions does (update_positions.cl:9-12); nullptr stops it.  The
  // current state is written at once.
  void setSharingBuffer(void *devicePositions);

  // The underlying C-ABI handle (nullptr before init()).
  pbf_sim *handle() const { return mSim; }

 private:
  Float4 mSystemSizeMin;
  Float4 mSystemSizeMax;
  const unsigned mNumParticles;
  pbf_params mParams;
  const std::vector<Particle> &mParticles;
  std::vector<Float4> mPositions;
  std::vector<Float4> mVelocities;
  pbf_sim *mSim;
  float mWaveGenerator;
  void *mSharingBuffer;
};
