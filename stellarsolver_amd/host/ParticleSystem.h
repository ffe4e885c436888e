// Abstract particle system (the interface of ParticleSystem.h:7-23 in the
// reference, which its BarnesHutParticleSystem never derived from; here the
// Barnes-Hut system implements it).
#ifndef STELLAR_PARTICLE_SYSTEM_H
#define STELLAR_PARTICLE_SYSTEM_H

#include "SimulationParameters.h"

class ParticleSystem {
 public:
  SimulationParameters parameters;

  ParticleSystem(const SimulationParameters p, const int n) : parameters(p) { (void)n; }
  virtual ~ParticleSystem() {}

  virtual int getNumParticles() = 0;
  virtual void update() = 0;
  virtual void reset() = 0;
  virtual float* getOutputBuffer() = 0;
};

#endif
