// Options of the drop-in host side that the reference does not have as
// parameters: it hard-codes theta = 0.5 and eps2 = 0.025 (kernels.cu:8-9),
// walks in warps of 32 and always starts from its own three models.  Every
// member has a default, so BarnesHutParticleSystem(parameters, n) -- the
// reference's constructor -- behaves as the reference configures its step.
#ifndef STELLAR_OPTIONS_H
#define STELLAR_OPTIONS_H

#include <cstdint>

struct StellarOptions {
  float theta = 0.5f;      // opening parameter (kernels.cu:9)
  float eps2 = 0.025f;     // softening squared (kernels.cu:8)
  int32_t group = 64;      // walk-group size: 64 (one wavefront) or 32 (the reference's warp)
  int32_t workload = -1;   // -1: the reference's models; 0, 1, 2, 4: seeded workloads
  bool seed_given = false;
  uint64_t seed = 0;       // seed of a seeded workload when seed_given
  bool quadrupole = false; // add cell quadrupole terms
  bool textbook = false;   // opening test w/d <= theta on the unsoftened distance
  bool hilbert = false;    // bodies and walk groups along a Hilbert curve of the tree
};

#endif
