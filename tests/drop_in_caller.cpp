// A caller written the way the reference's main.cpp:78-97 drives its
// BarnesHutParticleSystem: it declares SimulationParameters on the stack,
// sets the nine fields the reference has, and constructs the system with
// (parameters, n).  Compiled by tests/test_drop_in.py against the drop-in
// headers; with "--step" it also builds the system and runs one update().
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <exception>

#include "BarnesHutParticleSystem.h"
#include "SimulationParameters.h"

// Leaves non-zero bytes where the next frame's locals will live, so that a
// field the caller does not set would not read as zero by luck.
__attribute__((noinline)) static void dirty_stack() {
  volatile unsigned char junk[8192];
  for (size_t i = 0; i < sizeof(junk); ++i) junk[i] = 0xAB;
}

__attribute__((noinline)) static int run(bool step) {
  int numbodies = 4096;
  SimulationParameters parameters;
  parameters.model = disk_model;
  parameters.opengl = false;
  parameters.debug = false;
  parameters.benchmark = false;
  parameters.fullscreen = false;
  parameters.iterations = 1;
  parameters.timestep = 0.001f;
  parameters.gravity = 1.0f;
  parameters.dampening = 1.0f;
  std::printf("layout %zu %zu %zu %zu\n", sizeof(SimulationParameters),
              offsetof(SimulationParameters, fullscreen), offsetof(SimulationParameters, iterations),
              offsetof(SimulationParameters, dampening));
  const stellar_params sp = BarnesHutParticleSystem::contextParameters(parameters);
  std::printf("params %.9g %.9g %d %d %.9g\n", sp.theta, sp.eps2, sp.group, sp.flags, sp.timestep);
  if (!step) return 0;
  try {
    BarnesHutParticleSystem particles(parameters, numbodies);
    particles.reset();
    particles.update();
    stellar_params got;
    if (stellar_get_params(particles.context(), &got) != STELLAR_OK) return 2;
    const float* out = particles.getOutputBuffer();
    std::printf("context %.9g %.9g %d %d\n", got.theta, got.eps2, got.group, got.flags);
    std::printf("output %d %.9g %.9g\n", particles.getNumParticles(), out[0], out[1]);
  } catch (const std::exception& e) {
    std::printf("error %s\n", e.what());
    return 3;
  }
  return 0;
}

int main(int argc, char** argv) {
  dirty_stack();
  return run(argc > 1 && std::strcmp(argv[1], "--step") == 0);
}
