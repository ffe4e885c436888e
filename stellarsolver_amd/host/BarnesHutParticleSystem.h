// Drop-in for the reference's BarnesHutParticleSystem
// (BarnesHutParticleSystem.h:9-78): same constructor, getNumParticles(),
// update(), reset() and getOutputBuffer(), over the libstellar_hip C ABI
// instead of raw device buffers and CUDA launch wrappers.
//
// Differences a caller can see:
//  * errors throw std::runtime_error with the library's message (the
//    reference prints and then waits on stdin, BarnesHutParticleSystem.cpp:20);
//  * copies share nothing: the copy constructor and assignment copy the
//    simulation state (the reference's reallocate but leave the new buffers
//    uninitialised, :87-246).
#ifndef STELLAR_BARNES_HUT_PARTICLE_SYSTEM_H
#define STELLAR_BARNES_HUT_PARTICLE_SYSTEM_H

#include <string>
#include <vector>

#include "ParticleSystem.h"
#include "stellar.h"

class BarnesHutParticleSystem : public ParticleSystem {
 public:
  BarnesHutParticleSystem(const SimulationParameters p, const int n);
  BarnesHutParticleSystem(const BarnesHutParticleSystem& system);
  BarnesHutParticleSystem& operator=(const BarnesHutParticleSystem& system);
  ~BarnesHutParticleSystem() override;

  int getNumParticles() override;
  void update() override;
  void reset() override;
  // Interleaved (x0, y0, x1, y1, ...) positions after the last update, body-id
  // order; valid until the next update() or destruction.
  float* getOutputBuffer() override;

  // Milliseconds of the last update() (the value -benchmark prints).
  float lastStepMilliseconds() const { return last_ms_; }
  stellar_ctx* context() { return ctx_; }

 private:
  void create();
  void check(int status, const char* what) const;
  void runDebugTests();

  int numParticles_;
  int step_;
  stellar_ctx* ctx_;
  std::vector<float> output_;
  float last_ms_;
};

#endif
