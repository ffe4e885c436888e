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
#include "StellarOptions.h"
#include "stellar.h"

class BarnesHutParticleSystem : public ParticleSystem {
 public:
  // The reference's constructor: options at their defaults (theta 0.5,
  // eps2 0.025 as the reference hard-codes them; walk group 64).
  BarnesHutParticleSystem(const SimulationParameters p, const int n);
  BarnesHutParticleSystem(const SimulationParameters p, const int n, const StellarOptions& opt);
  BarnesHutParticleSystem(const BarnesHutParticleSystem& system);
  BarnesHutParticleSystem& operator=(const BarnesHutParticleSystem& system);
  ~BarnesHutParticleSystem() override;

  int getNumParticles() override;
  void update() override;
  void reset() override;
  // Interleaved (x0, y0, x1, y1, ...) positions after the last update, body-id
  // order; valid until the next update() or destruction.
  float* getOutputBuffer() override;

  // Headless form of the reference's OpenGL drawing of the output buffer
  // (StellarSolverVisualizer.cpp:208-253): an 8-bit grey image, row 0 at the
  // top, of the current positions through view = {left, right, bottom, top}
  // (the reference: -25, 25, -25, 25 into 1000 x 1000), each body adding
  // white at alpha 0.1 (stellar_render).
  const std::vector<uint8_t>& renderFrame(const float view[4], int width, int height);

  // Milliseconds of the last update() (the value -benchmark prints).
  float lastStepMilliseconds() const { return last_ms_; }
  stellar_ctx* context() { return ctx_; }
  const StellarOptions& options() const { return options_; }
  // The library parameters a system built from (p, opt) runs with.
  static stellar_params contextParameters(const SimulationParameters& p,
                                          const StellarOptions& opt = StellarOptions());

 private:
  void create();
  void check(int status, const char* what) const;
  void runDebugTests();

  StellarOptions options_;
  int numParticles_;
  int step_;
  stellar_ctx* ctx_;
  std::vector<float> output_;
  std::vector<uint8_t> frame_;
  float last_ms_;
};

#endif
