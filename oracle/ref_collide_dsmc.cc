// ref_collide_dsmc.cc -- C entry point around the reference's own DSMC cell lists and
// collide, cellTracking and collisions
// (thesisCode/dsmcCollisions/cellTrackerAdvanced/src/main.cc:269-350), compiled from where
// they lie together with that variant's particle.cc and grid.cc.  The reference's main()
// in the same file calls its CUDA entry point and FFTW; built with hidden visibility and
// section garbage collection it is unreferenced and dropped by the linker, so nothing else
// is needed.  Test infrastructure only.
#include <cstdint>
#include <memory>
#include <random>
#include <vector>

#include "grid.h"
#include "params.h"
#include "particle.h"

void cellTracking(std::vector<Particle> particles, Grid& grid, std::shared_ptr<Params> p);
void collisions(std::mt19937& gen, std::vector<Particle>& particles, std::shared_ptr<Params> p,
                Grid& grid);

extern "C" __attribute__((visibility("default")))
// nsteps passes of the reference's per-step DSMC work (resetGrid, cellTracking,
// collisions; positions fixed) over n particles on an nx x ny grid of cells dx x dy, with
// one std::mt19937 gen{seed} for the whole run.
void ref_dsmc_collide(int64_t n, const double* x, const double* y, double* vx, double* vy,
                      int nx, int ny, double dx, double dy, uint32_t seed, int nsteps) {
  auto p = std::make_shared<Params>();
  p->nx = nx;
  p->ny = ny;
  p->dx = dx;
  p->dy = dy;
  p->area = dx * dy;
  p->numParticles = (int)n;
  std::vector<Particle> parts;
  parts.reserve(n);
  for (int64_t i = 0; i < n; ++i) {
    parts.emplace_back(Point{x[i], y[i]}, p);
    parts.back().setVelocity(Point{vx[i], vy[i]});
  }
  Grid g(nx, ny, (int)n, p);
  std::mt19937 gen{seed};
  for (int s = 0; s < nsteps; ++s) {
    g.resetGrid();
    cellTracking(parts, g, p);
    collisions(gen, parts, p, g);
  }
  for (int64_t i = 0; i < n; ++i) {
    const Point v = parts[i].getVelocity();
    vx[i] = v.x;
    vy[i] = v.y;
  }
}
