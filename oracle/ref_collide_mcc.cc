// ref_collide_mcc.cc -- C entry point around the reference's own MCC collide,
// Particle::collision (thesisCode/mccCollisions/src/particle.cc:112-127), compiled from
// where it lies.  Test infrastructure only: tests/ feed the reference and the GPU operator
// the same std::mt19937 stream and compare the velocities bit for bit.
#include <cstdint>
#include <memory>
#include <random>
#include <vector>

#include "grid.h"
#include "params.h"
#include "particle.h"

extern "C" __attribute__((visibility("default")))
// nsteps passes of the reference's collision loop (main.cc:181-183: every particle in
// order, one std::mt19937 gen{seed} for the whole run) over n particles.
void ref_mcc_collide(int64_t n, double* vx, double* vy, uint32_t seed, int nsteps) {
  auto p = std::make_shared<Params>();
  p->nx = p->ny = 4;
  p->dx = p->dy = 1.0;
  p->area = 1.0;
  p->numParticles = (int)n;
  std::vector<Particle> parts;
  parts.reserve(n);
  for (int64_t i = 0; i < n; ++i) {
    parts.emplace_back(Point{0.5, 0.5}, p);
    parts.back().setVelocity(Point{vx[i], vy[i]});
  }
  std::mt19937 gen{seed};
  for (int s = 0; s < nsteps; ++s)
    for (auto& q : parts) q.collision(gen);
  for (int64_t i = 0; i < n; ++i) {
    const Point v = parts[i].getVelocity();
    vx[i] = v.x;
    vy[i] = v.y;
  }
}
