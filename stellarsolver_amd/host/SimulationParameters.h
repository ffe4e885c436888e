// Simulation parameters of the drop-in host side: the reference's
// SimulationParameters and Model (SimulationParameters.h:6-25), field for
// field, so that a caller written for the reference (main.cpp:78-97, which
// declares the struct and sets these nine fields) compiles and runs
// unchanged.  The knobs the reference hard-codes (theta, softening, walk
// group) live in StellarOptions (StellarOptions.h), which has defaults.
#ifndef STELLAR_SIMULATION_PARAMETERS_H
#define STELLAR_SIMULATION_PARAMETERS_H

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
} SimulationParameters;

#endif
