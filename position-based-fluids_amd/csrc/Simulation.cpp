// Simulation.cpp - C++ facade over the C ABI; errors from the ABI become
// std::runtime_error as the reference's cl::Error / runtime_error did.
#include "Simulation.hpp"

#include <stdexcept>
#include <string>

namespace {
void check(int code, const char *what) {
  if (code != PBF_OK)
    throw std::runtime_error(std::string(what) + ": " + pbf_last_error());
}
}  // namespace

Simulation::Simulation(const ConfigParameters &parameters,
                       const std::vector<Particle> &particles)
    : mNumParticles((unsigned)particles.size()),
      mParams(parameters.toPbfParams()),
      mParticles(particles),
      mSim(nullptr),
      mWaveGenerator(0.0f),
      mSharingBuffer(nullptr) {
  mSystemSizeMin.s[0] = parameters.xMin;
  mSystemSizeMin.s[1] = parameters.yMin;
  mSystemSizeMin.s[2] = parameters.zMin;
  mSystemSizeMin.s[3] = 0.0f;
  mSystemSizeMax.s[0] = parameters.xMax;
  mSystemSizeMax.s[1] = parameters.yMax;
  mSystemSizeMax.s[2] = parameters.zMax;
  mSystemSizeMax.s[3] = 0.0f;
}

Simulation::~Simulation() { pbf_destroy(mSim); }

void Simulation::init(void) {
  if (mSim) pbf_destroy(mSim);
  mSim = nullptr;
  static_assert(sizeof(Particle) == sizeof(pbf_particle), "layout");
  check(pbf_create(&mParams,
                   reinterpret_cast<const pbf_particle *>(mParticles.data()),
                   mNumParticles, &mSim),
        "Simulation::init");
  if (mWaveGenerator != 0.0f)
    check(pbf_set_wave_generator(mSim, mWaveGenerator), "setWaveGenerator");
  if (mSharingBuffer)
    check(pbf_positions_to_device(mSim, 0, mSharingBuffer), "Simulation::init");
}

void Simulation::initCells(void) {}

void Simulation::step(void) {
  if (!mSim) throw std::runtime_error("Simulation::step before init()");
  check(pbf_step(mSim, 1), "Simulation::step");
  check(pbf_synchronize(mSim), "Simulation::step");
  if (mSharingBuffer)
    check(pbf_positions_to_device(mSim, 0, mSharingBuffer), "Simulation::step");
}

void Simulation::dumpData(Float4 *(&positions), Float4 *(&velocities)) {
  if (!mSim) throw std::runtime_error("Simulation::dumpData before init()");
  mPositions.resize(mNumParticles);
  mVelocities.resize(mNumParticles);
  check(pbf_dump(mSim, reinterpret_cast<float *>(mPositions.data()),
                 reinterpret_cast<float *>(mVelocities.data())),
        "Simulation::dumpData");
  positions = mPositions.data();
  velocities = mVelocities.data();
}

void Simulation::setWaveGenerator(const double value) {
  mWaveGenerator = (float)value;
  if (mSim)
    check(pbf_set_wave_generator(mSim, mWaveGenerator), "setWaveGenerator");
}

void Simulation::setSharingBuffer(void *devicePositions) {
  mSharingBuffer = devicePositions;
  if (mSim && mSharingBuffer)
    check(pbf_positions_to_device(mSim, 0, mSharingBuffer),
          "setSharingBuffer");
}
