// ref_harness.cc -- C entry points around the reference's own Grid and Particle classes
// (thesisCode/collisionless/double/src/{grid.cc,particle.cc}, compiled from where they lie).
// Test infrastructure only: lets tests/ compare the oracle's CIC deposit, gather and push
// with the reference code through the documented bridges (reflected weights, lagged force).
#include <memory>
#include <vector>

#include "grid.h"
#include "params.h"
#include "particle.h"

extern "C" {

// Reference chargeAssignment (grid.cc:46-76) on n particles; rho_pad receives the
// (nx+2)*(ny+2) ghost-padded array as the reference leaves it.
void ref_charge_assignment(int n, const double* x, const double* y, int xdim, int ydim, int nx,
                           int ny, double* rho_pad) {
  auto p = std::make_shared<Params>();
  p->xDim = xdim; p->yDim = ydim; p->nx = nx; p->ny = ny;
  p->dx = (double)xdim / nx; p->dy = (double)ydim / ny; p->area = p->dx * p->dy;
  p->numParticles = n;
  Grid g(nx, ny, n, p);
  std::vector<Particle> parts;
  parts.reserve(n);
  for (int i = 0; i < n; ++i) parts.emplace_back(Point{x[i], y[i]}, p);
  g.chargeAssignment(parts);
  for (size_t i = 0; i < g.chargeDensity.size(); ++i) rho_pad[i] = g.chargeDensity[i];
}

// Reference forceCalculation (particle.cc:37-60) with a caller-supplied padded field
// (ex_pad/ey_pad are (nx+2)*(ny+2) arrays indexed as the reference indexes them).
void ref_force(int n, const double* x, const double* y, int xdim, int ydim, int nx, int ny,
               const double* ex_pad, const double* ey_pad, double* fx, double* fy) {
  auto p = std::make_shared<Params>();
  p->xDim = xdim; p->yDim = ydim; p->nx = nx; p->ny = ny;
  p->dx = (double)xdim / nx; p->dy = (double)ydim / ny; p->area = p->dx * p->dy;
  p->numParticles = n;
  Grid g(nx, ny, n, p);
  for (size_t i = 0; i < g.electricField.size(); ++i)
    g.electricField[i] = Point{ex_pad[i], ey_pad[i]};
  for (int i = 0; i < n; ++i) {
    Particle q(Point{x[i], y[i]}, p);
    q.setVelocity(Point{0.0, 0.0});  // the constructor leaves the velocity unset
    q.forceCalculation(g, p);
    q.positionUpdate(p);  // velocity 0, forceOld 0: only v = dt/2 * force changes
    Point v = q.getVelocity();
    fx[i] = v.x / (0.5 * p->dt);
    fy[i] = v.y / (0.5 * p->dt);
  }
}

// Reference positionUpdate (particle.cc:69-106): one call from a fresh particle
// (forceOld = 0, force = 0): x += dt*v with the reference's wrap.
void ref_drift(int n, double* x, double* y, const double* vx, const double* vy, int xdim,
               int ydim, int nx, int ny, double dt) {
  auto p = std::make_shared<Params>();
  p->xDim = xdim; p->yDim = ydim; p->nx = nx; p->ny = ny;
  p->dx = (double)xdim / nx; p->dy = (double)ydim / ny; p->area = p->dx * p->dy;
  p->dt = dt;
  for (int i = 0; i < n; ++i) {
    Particle q(Point{x[i], y[i]}, p);
    q.setVelocity(Point{vx[i], vy[i]});
    q.positionUpdate(p);
    Point r = q.getPosition();
    x[i] = r.x;
    y[i] = r.y;
  }
}

}  // extern "C"
