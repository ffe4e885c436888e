// Simulation parameters of the drop-in host side.
//
// The first nine fields are the reference's SimulationParameters
// (SimulationParameters.h:6-25), with the same meaning; the rest are the
// knobs the reference hard-codes (N, theta, softening, launch shape) or does
// not have.
#ifndef STELLAR_SIMULATION_PARAMETERS_H
#define STELLAR_SIMULATION_PARAMETERS_H

#include <cstdint>

typedef enum Model {
  disk_model,
  colliding_disk_model,
  plummer_model
} Model;

typedef struct SimulationParameters {
  Model model;
  bool opengl;
  bool debug;
  bool benchmark;
  bool fullscreen;
  float iterations;
  float timestep;
  float gravity;
  float dampening;
  // ---- not in the reference ----
  float theta;       // opening parameter (reference: 0.5, kernels.cu:9)
  float eps2;        // softening squared (reference: 0.025, kernels.cu:8)
  int32_t group;     // walk-group size: 64 (native) or 32 (reference warp)
  int32_t workload;  // -1: the reference's own models; 0,1,2,4: seeded workloads
  uint64_t seed;     // seed of a seeded workload (default: the workload's)
  bool seed_given;
  bool quadrupole;   // add cell quadrupole terms
  bool textbook;     // opening test w/d <= theta
} SimulationParameters;

#endif
