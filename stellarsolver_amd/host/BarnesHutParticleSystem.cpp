// BarnesHutParticleSystem over libstellar_hip (see BarnesHutParticleSystem.h).
#include "BarnesHutParticleSystem.h"

#include <chrono>
#include <cstdint>
#include <iostream>
#include <stdexcept>
#include <vector>

stellar_params BarnesHutParticleSystem::contextParameters(const SimulationParameters& p,
                                                          const StellarOptions& opt) {
  stellar_params s;
  stellar_default_params(&s);
  s.theta = opt.theta;
  s.eps2 = opt.eps2;
  s.timestep = p.timestep;
  s.dampening = p.dampening;
  s.gravity = p.gravity;
  s.group = opt.group;
  s.flags = (opt.quadrupole ? STELLAR_FLAG_QUADRUPOLE : 0) |
            (opt.textbook ? STELLAR_FLAG_TEXTBOOK_OPENING : 0) |
            (opt.hilbert ? STELLAR_FLAG_HILBERT_ORDER : 0);
  return s;
}

BarnesHutParticleSystem::BarnesHutParticleSystem(const SimulationParameters p, const int n)
    : BarnesHutParticleSystem(p, n, StellarOptions()) {}

BarnesHutParticleSystem::BarnesHutParticleSystem(const SimulationParameters p, const int n,
                                                 const StellarOptions& opt)
    : ParticleSystem(p, n), options_(opt), numParticles_(n), step_(0), ctx_(nullptr),
      last_ms_(0.f) {
  create();
}

BarnesHutParticleSystem::BarnesHutParticleSystem(const BarnesHutParticleSystem& system)
    : ParticleSystem(system.parameters, system.numParticles_),
      options_(system.options_),
      numParticles_(system.numParticles_),
      step_(system.step_),
      ctx_(nullptr),
      last_ms_(system.last_ms_) {
  create();
  const size_t n = (size_t)numParticles_;
  std::vector<float> m(n), x(n), y(n), vx(n), vy(n);
  check(stellar_download(system.ctx_, m.data(), x.data(), y.data(), vx.data(), vy.data(),
                         nullptr, nullptr),
        "download");
  check(stellar_upload(ctx_, m.data(), x.data(), y.data(), vx.data(), vy.data()), "upload");
  output_ = system.output_;
}

BarnesHutParticleSystem& BarnesHutParticleSystem::operator=(const BarnesHutParticleSystem& system) {
  if (this != &system) {
    BarnesHutParticleSystem tmp(system);
    std::swap(parameters, tmp.parameters);
    std::swap(options_, tmp.options_);
    std::swap(numParticles_, tmp.numParticles_);
    std::swap(step_, tmp.step_);
    std::swap(ctx_, tmp.ctx_);
    std::swap(output_, tmp.output_);
    std::swap(last_ms_, tmp.last_ms_);
  }
  return *this;
}

BarnesHutParticleSystem::~BarnesHutParticleSystem() {
  if (ctx_) stellar_destroy(ctx_);
}

void BarnesHutParticleSystem::create() {
  const stellar_params sp = contextParameters(parameters, options_);
  const int rc = stellar_create(&sp, numParticles_, 0, &ctx_);
  if (rc != STELLAR_OK) {
    ctx_ = nullptr;
    throw std::runtime_error(std::string("stellar_create: ") + stellar_status_string(rc));
  }
}

void BarnesHutParticleSystem::check(int status, const char* what) const {
  if (status != STELLAR_OK)
    throw std::runtime_error(std::string(what) + ": " + stellar_last_error(ctx_));
}

int BarnesHutParticleSystem::getNumParticles() { return numParticles_; }

// BarnesHutParticleSystem::update (BarnesHutParticleSystem.cpp:301-347): one
// synchronous step; -benchmark prints the step time, -debug runs the tree
// checks of debug.cpp on the tree just built.
void BarnesHutParticleSystem::update() {
  const auto t0 = std::chrono::steady_clock::now();
  check(stellar_step(ctx_), "stellar_step");
  const auto t1 = std::chrono::steady_clock::now();
  last_ms_ = std::chrono::duration<float, std::milli>(t1 - t0).count();
  if (parameters.benchmark)
    std::cout << "Timestep: " << step_ << "    " << "Elapsed time: " << last_ms_ << std::endl;
  if (parameters.opengl) getOutputBuffer();
  if (parameters.debug) runDebugTests();
  ++step_;
}

// BarnesHutParticleSystem::reset (BarnesHutParticleSystem.cpp:350-372).
void BarnesHutParticleSystem::reset() {
  stellar_ic_params ic;
  if (options_.workload < 0) {
    stellar_ic_defaults(0, &ic);
    ic.model = parameters.model == plummer_model          ? STELLAR_IC_REF_PLUMMER
               : parameters.model == colliding_disk_model ? STELLAR_IC_REF_COLLIDING
                                                          : STELLAR_IC_REF_DISK;
  } else {
    if (stellar_ic_defaults(options_.workload, &ic) != STELLAR_OK)
      throw std::runtime_error("unknown workload");
    if (options_.seed_given) ic.seed = options_.seed;
  }
  ic.gravity = parameters.gravity;
  const size_t n = (size_t)numParticles_;
  std::vector<float> m(n), x(n), y(n), vx(n), vy(n);
  check(stellar_generate_ic(&ic, numParticles_, m.data(), x.data(), y.data(), vx.data(),
                            vy.data()),
        "stellar_generate_ic");
  check(stellar_upload(ctx_, m.data(), x.data(), y.data(), vx.data(), vy.data()),
        "stellar_upload");
  step_ = 0;
}

const std::vector<uint8_t>& BarnesHutParticleSystem::renderFrame(const float view[4], int width,
                                                                int height) {
  frame_.resize((size_t)width * (size_t)height);
  check(stellar_render(ctx_, view, width, height, nullptr, frame_.data()), "stellar_render");
  return frame_;
}

float* BarnesHutParticleSystem::getOutputBuffer() {
  output_.resize(2 * (size_t)numParticles_);
  check(stellar_download_positions(ctx_, output_.data()), "stellar_download_positions");
  return output_.data();
}

// The checks of DEBUG_RUN_TESTS (debug.cpp:6-178) on the tree of the last
// step, laid out as the reference's pool: cells at n+c, root slots in
// child[0..3], m = 2n+12000 nodes.
void BarnesHutParticleSystem::runDebugTests() {
  const int64_t n = numParticles_;
  int64_t ncells = 0;
  check(stellar_num_cells(ctx_, &ncells), "stellar_num_cells");
  const int64_t cap = ncells > 0 ? ncells : 1;
  std::vector<int32_t> root(4), child(4 * cap), count(cap), start(cap), level(cap), sorted(n);
  check(stellar_export_tree(ctx_, root.data(), child.data(), count.data(), start.data(),
                            level.data(), sorted.data(), nullptr, nullptr, nullptr, cap,
                            &ncells),
        "stellar_export_tree");
  float bbox[4];
  check(stellar_get_bbox(ctx_, bbox), "stellar_get_bbox");
  const int64_t m = std::max<int64_t>(2 * n + 12000, n + ncells);
  std::vector<int64_t> pchild(4 * m, -1), pcount(m, 0), psorted(m, 0);
  for (int64_t i = 0; i < n; ++i) pcount[i] = 1;
  for (int k = 0; k < 4; ++k) pchild[k] = root[k];
  for (int64_t c = 0; c < ncells; ++c) {
    pcount[n + c] = count[c];
    for (int k = 0; k < 4; ++k) pchild[4 * (n + c) + k] = child[4 * c + k];
  }
  for (int64_t i = 0; i < n; ++i) psorted[i] = sorted[i];
  auto report = [](const char* name, bool pass) {
    std::cout << name << " RESULT = " << (pass ? "PASS: " : "FAIL: ") << std::endl;
  };
  bool ok = true;  // count test 1
  for (int64_t i = n; i < m; ++i) {
    int64_t c = pcount[i];
    for (int j = 0; j < 4; ++j) {
      const int64_t ch = pchild[4 * i + j];
      if (ch != -1 && ch >= 0 && ch < m) c -= pcount[ch];
    }
    if (c != 0) ok = false;
  }
  report("COUNT ARRAY TEST 1", ok);
  int64_t sum = 0;  // child tests 1-3
  bool below_m = true, above_m1 = true;
  for (int64_t j = 0; j < 4 * m; ++j) {
    if (pchild[j] >= 0 && pchild[j] < n) sum += pchild[j];
    if (pchild[j] >= m) below_m = false;
    if (pchild[j] < -1) above_m1 = false;
  }
  report("CHILD ARRAY TEST 1", sum == n * (n - 1) / 2);
  report("CHILD ARRAY TEST 2", below_m);
  report("CHILD ARRAY TEST 3", above_m1);
  int64_t mn = 2 * n;  // child test 4
  for (int k = 0; k < 4; ++k) mn = std::min(mn, pchild[k]);
  ok = true;
  for (int64_t j = 4; j < 4 * (mn - 1); ++j)
    if (pchild[j] != -1) ok = false;
  report("CHILD ARRAY TEST 4", ok);
  sum = 0;  // sorted tests 1-2
  for (int64_t i = 0; i < m; ++i) sum += psorted[i];
  report("SORTED ARRAY TEST 1", sum == n * (n - 1) / 2);
  std::vector<char> seen(n, 0);
  for (int64_t i = 0; i < n; ++i)
    if (psorted[i] >= 0 && psorted[i] < n) seen[psorted[i]] = 1;
  ok = true;
  for (int64_t i = 0; i < n; ++i) ok = ok && seen[i];
  report("SORTED ARRAY TEST 2", ok);
  report("BOUNDING BOX TEST 1", bbox[0] <= bbox[1] && bbox[2] <= bbox[3]);
}
