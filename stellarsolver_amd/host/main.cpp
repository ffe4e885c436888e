// stellar_app: the reference's command line (main.cpp:75-103) and headless
// simulation loop (StellarSolverVisualizer::runSimulation,
// StellarSolverVisualizer.cpp:199-206) over BarnesHutParticleSystem.
//
// Reference flags: -disk (default), -plummer, -colliding-disks, -opengl,
// -debug, -benchmark, -fullscreen, -iterations=, -timestep=, -gravity=,
// -dampening= (and -barnes-hut, which the README shows and which selects
// nothing).  Added: -numbodies= (reference: 512*256), -theta=, -eps=,
// -group=, -workload= (seeded workload 0, 1, 2 or 4), -seed=, -quad,
// -textbook, -hilbert; -frames=DIR, -frame-every=K, -view=S, -resolution=R.
//
// -opengl: the reference opens a window and draws every body after every
// step (StellarSolverVisualizer.cpp:208-253).  Headless here: every step is
// rendered on the GPU into the reference's view (glm::ortho(-S, S, -S, S),
// S = 25, into R x R pixels, R = 1000) as an 8-bit grey image, and with
// -frames=DIR every K-th image is written as DIR/frame_NNNNN.pgm (-frames
// alone implies -opengl).  As in the reference, a flag's first character is skipped
// when matching, so "-x" and "--x"-style prefixes of one character both work.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "BarnesHutParticleSystem.h"
#include "SimulationParameters.h"
#include "StellarOptions.h"

namespace {

bool checkCmdLineFlag(const int argc, char** argv, const char* string) {
  for (int i = 1; i < argc; i++)
    if (argv[i][0] != '\0' && std::strcmp(&argv[i][1], string) == 0) return true;
  return false;
}

// Value text of "-name=value", or nullptr.
const char* cmdLineValue(const int argc, char** argv, const char* name) {
  const size_t len = std::strlen(name);
  for (int i = 1; i < argc; i++) {
    if (argv[i][0] == '\0') continue;
    const char* str = &argv[i][1];
    const char* eqs = std::strchr(str, '=');
    if (eqs && (size_t)(eqs - str) == len && std::strncmp(str, name, len) == 0) return eqs + 1;
  }
  return nullptr;
}

int getCmdLineParameterInt(const int argc, char** argv, const char* s, const int def) {
  const char* v = cmdLineValue(argc, argv, s);
  return v ? std::atoi(v) : def;
}

float getCmdLineParameterFloat(const int argc, char** argv, const char* s, const float def) {
  const char* v = cmdLineValue(argc, argv, s);
  return v ? (float)std::atof(v) : def;
}

// Binary PGM (P5) of a grey image, row 0 at the top.
void writePgm(const char* path, const std::vector<uint8_t>& img, int w, int h) {
  FILE* f = std::fopen(path, "wb");
  if (!f) throw std::runtime_error(std::string("cannot write ") + path);
  std::fprintf(f, "P5\n%d %d\n255\n", w, h);
  const size_t wrote = std::fwrite(img.data(), 1, img.size(), f);
  std::fclose(f);
  if (wrote != img.size()) throw std::runtime_error(std::string("short write to ") + path);
}

}  // namespace

int main(int argc, char** argv) {
  int numbodies = getCmdLineParameterInt(argc, argv, "numbodies", 512 * 256);
  SimulationParameters parameters;
  if (checkCmdLineFlag(argc, argv, "plummer"))
    parameters.model = plummer_model;
  else if (checkCmdLineFlag(argc, argv, "colliding-disks"))
    parameters.model = colliding_disk_model;
  else
    parameters.model = disk_model;
  parameters.opengl = checkCmdLineFlag(argc, argv, "opengl");
  parameters.debug = checkCmdLineFlag(argc, argv, "debug");
  parameters.benchmark = checkCmdLineFlag(argc, argv, "benchmark");
  parameters.fullscreen = checkCmdLineFlag(argc, argv, "fullscreen");
  parameters.iterations = (float)getCmdLineParameterInt(argc, argv, "iterations", 50);
  parameters.timestep = getCmdLineParameterFloat(argc, argv, "timestep", 0.001f);
  parameters.gravity = getCmdLineParameterFloat(argc, argv, "gravity", 1.0f);
  parameters.dampening = getCmdLineParameterFloat(argc, argv, "dampening", 1.0f);
  StellarOptions options;
  options.theta = getCmdLineParameterFloat(argc, argv, "theta", options.theta);
  const char* eps = cmdLineValue(argc, argv, "eps");
  if (eps) options.eps2 = (float)(std::atof(eps) * std::atof(eps));
  options.group = getCmdLineParameterInt(argc, argv, "group", options.group);
  options.workload = getCmdLineParameterInt(argc, argv, "workload", options.workload);
  const char* seed = cmdLineValue(argc, argv, "seed");
  options.seed_given = seed != nullptr;
  if (seed) options.seed = (uint64_t)std::strtoull(seed, nullptr, 10);
  options.quadrupole = checkCmdLineFlag(argc, argv, "quad");
  options.textbook = checkCmdLineFlag(argc, argv, "textbook");
  options.hilbert = checkCmdLineFlag(argc, argv, "hilbert");
  const char* frames = cmdLineValue(argc, argv, "frames");
  if (frames) parameters.opengl = true;
  const int frameEvery = std::max(1, getCmdLineParameterInt(argc, argv, "frame-every", 1));
  const float viewHalf = getCmdLineParameterFloat(argc, argv, "view", 25.0f);
  const int resolution = getCmdLineParameterInt(argc, argv, "resolution", 1000);
  const float view[4] = {-viewHalf, viewHalf, -viewHalf, viewHalf};
  try {
    BarnesHutParticleSystem particles(parameters, numbodies, options);
    particles.reset();
    double total_ms = 0.0;
    for (int i = 0; i < (int)parameters.iterations; i++) {
      particles.update();
      total_ms += particles.lastStepMilliseconds();
      if (parameters.opengl) {
        const std::vector<uint8_t>& img = particles.renderFrame(view, resolution, resolution);
        if (frames && i % frameEvery == 0) {
          char path[4096];
          std::snprintf(path, sizeof path, "%s/frame_%05d.pgm", frames, i);
          writePgm(path, img, resolution, resolution);
        }
      }
    }
    if (parameters.benchmark && parameters.iterations > 0)
      std::cout << "Bodies: " << numbodies << "    Steps: " << (int)parameters.iterations
                << "    Mean step: " << total_ms / parameters.iterations << " ms    "
                << "Body-steps/s: "
                << (double)numbodies * parameters.iterations / (total_ms * 1e-3) << std::endl;
  } catch (const std::exception& e) {
    std::cerr << "stellar_app: " << e.what() << std::endl;
    return 1;
  }
  return 0;
}
