// pic_main -- command-line driver of libpicamd.so.
//
// Accepts the reference driver's flags (thesisCode/collisionless/double/src/main.cc:41-72):
//   -x XDIM -y YDIM   domain size (Debye lengths)       -l N      grid points per axis
//   -n NPX  -m NPY    particle lattice                  -t TMAX   simulated time
//   -f FILE           diagnostics file (CSV, one row per step; the reference only
//                     truncated it)                     -s BLOCK  accepted, ignored (the
//                     autotuner picks tile sizes)       -c        accepted: there is no CPU
//                     path here                         -g        skip the GPU run
// and adds:
//   -P PRESET  reference (default) | landau | twostream | mcc
//   -d DT      time step (default: the preset's; the reference's Params default 1e-7)
//   -N STEPS   number of steps (overrides -t)           -p f32|f64  particle precision
//   -r SEED    Philox key of the loaders and collisions -a        autotune before timing
//   -q PPC     particles per cell of the landau / twostream / mcc presets
//   -o FILE    final particles (CSV x,y,vx,vy in input order, synchronous velocities): the
//              positions the reference returns from picFloatGpu (xResult, yResult)
//   -h         this text
// The reference preset loads the reference's lattice and Maxwellian stream
// (pic_load_reference); the others are the configured workloads of the engine.
#include <getopt.h>
#include <sys/time.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "pic_capi.h"

namespace {

double now_s() {
  timeval t;
  gettimeofday(&t, nullptr);
  return t.tv_sec + 1e-6 * t.tv_usec;
}

void usage() {
  std::printf(
      "usage: pic_main [-x XDIM] [-y YDIM] [-l N] [-n NPX] [-m NPY] [-t TMAX] [-f FILE]\n"
      "                [-s BLOCK] [-c] [-g] [-P reference|landau|twostream|mcc] [-d DT]\n"
      "                [-N STEPS] [-p f32|f64] [-r SEED] [-q PPC] [-a] [-o FILE] [-h]\n");
}

bool check(int rc, const char* what) {
  if (rc == PIC_OK) return true;
  std::fprintf(stderr, "%s failed (%d): %s\n", what, rc, pic_last_error());
  return false;
}

bool pow2(int v) { return v >= 4 && (v & (v - 1)) == 0; }

}  // namespace

int main(int argc, char** argv) {
  // reference defaults (CLd/inc/params.h:21-47)
  int xdim = 10, ydim = 10, ngrid = 32, npx = 10, npy = 10;
  double tmax = 1e-4, dt = -1.0;
  std::string file, final_file, preset = "reference", prec = "f64";
  bool gpu = true, autotune = false;
  long long nsteps = -1;
  uint64_t seed = 1234;
  int ppc = 64;
  bool grid_set = false, dim_set = false;
  int opt;
  while ((opt = getopt(argc, argv, "x:y:l:n:m:t:f:s:cgP:d:N:p:r:q:ao:h")) != -1) {
    switch (opt) {
      case 'x': xdim = std::atoi(optarg); dim_set = true; break;
      case 'y': ydim = std::atoi(optarg); dim_set = true; break;
      case 'l': ngrid = std::atoi(optarg); grid_set = true; break;
      case 'n': npx = std::atoi(optarg); break;
      case 'm': npy = std::atoi(optarg); break;
      case 't': tmax = std::atof(optarg); break;
      case 'f': file = optarg; break;
      case 's': break;
      case 'c': break;
      case 'g': gpu = false; break;
      case 'P': preset = optarg; break;
      case 'd': dt = std::atof(optarg); break;
      case 'N': nsteps = std::atoll(optarg); break;
      case 'p': prec = optarg; break;
      case 'r': seed = std::strtoull(optarg, nullptr, 10); break;
      case 'q': ppc = std::atoi(optarg); break;
      case 'a': autotune = true; break;
      case 'o': final_file = optarg; break;
      case 'h': usage(); return 0;
      default: usage(); return 1;
    }
  }
  if (prec != "f32" && prec != "f64") {
    std::fprintf(stderr, "Error: -p takes f32 or f64\n");
    return 1;
  }
  pic_config cfg;
  pic_config_default(&cfg);
  cfg.precision = prec == "f64" ? PIC_FP64 : PIC_FP32;
  cfg.seed = seed;
  std::vector<double> x, y, vx, vy;
  const double four_pi = 4.0 * 3.14159265358979323846;
  if (preset == "reference") {
    if (npx < 1 || npy < 1 || xdim <= 0 || ydim <= 0 || !pow2(ngrid) || tmax < 0) {
      std::printf("Error: invalid input arguments\n");
      return 1;
    }
    cfg.nx = cfg.ny = ngrid;
    cfg.lx = xdim;
    cfg.ly = ydim;
    cfg.dt = dt > 0 ? dt : 1e-7;
    const size_t n = (size_t)npx * npy;
    x.resize(n); y.resize(n); vx.resize(n); vy.resize(n);
    if (!check(pic_load_reference(npx, npy, xdim, ydim, ngrid, ngrid, 1.0, 1.0, 0.0, 0.0, x.data(),
                                  y.data(), vx.data(), vy.data()),
               "pic_load_reference"))
      return 1;
  } else if (preset == "landau" || preset == "twostream" || preset == "mcc") {
    const int g = grid_set ? ngrid : (preset == "landau" ? 512 : (preset == "mcc" ? 256 : 64));
    if (!pow2(g) || ppc < 1) {
      std::printf("Error: invalid input arguments\n");
      return 1;
    }
    cfg.nx = cfg.ny = g;
    cfg.lx = cfg.ly = dim_set ? (double)xdim : (preset == "mcc" ? (double)g : four_pi);
    if (dim_set) cfg.ly = ydim;
    const int64_t n = pic_lattice_count(g, g, ppc);
    x.resize(n); y.resize(n); vx.resize(n); vy.resize(n);
    int rc;
    if (preset == "landau") {
      cfg.dt = dt > 0 ? dt : 0.05;
      rc = pic_load_landau(g, g, cfg.lx, cfg.ly, ppc, 1.0, 0.01, 1, seed, x.data(), y.data(),
                           vx.data(), vy.data());
    } else if (preset == "twostream") {
      cfg.dt = dt > 0 ? dt : 0.1;
      rc = pic_load_two_stream(g, g, cfg.lx, cfg.ly, ppc, 1.0, 0.01, 1e-3, 1, seed, x.data(),
                               y.data(), vx.data(), vy.data());
    } else {
      if (cfg.precision != PIC_FP32) {
        std::fprintf(stderr, "Error: the collision stages run on f32 particles\n");
        return 1;
      }
      if (!check(pic_mcc_preset(&cfg, 1e21, 2.0, 1e15), "pic_mcc_preset")) return 1;
      if (dt > 0) cfg.dt = dt;
      cfg.sort_every = 8;
      rc = pic_load_landau(g, g, cfg.lx, cfg.ly, ppc, 1.0, 0.0, 0, seed, x.data(), y.data(),
                           vx.data(), vy.data());
    }
    if (!check(rc, "loader")) return 1;
    if (nsteps < 0 && tmax == 1e-4) nsteps = preset == "mcc" ? 500 : (preset == "landau" ? 1000 : 200);
  } else {
    std::fprintf(stderr, "Error: unknown preset %s\n", preset.c_str());
    return 1;
  }
  const int64_t n = (int64_t)x.size();
  cfg.capacity = preset == "mcc" ? (int64_t)(1.25 * n) : n;
  if (nsteps < 0) nsteps = (long long)std::ceil(tmax / cfg.dt - 1e-9);
  if (nsteps > 0x7fffffff) {
    std::fprintf(stderr, "Error: too many steps\n");
    return 1;
  }
  std::printf("Preset: %s, grid %d x %d, domain %g x %g, %lld particles, dt %g, %lld steps, %s\n",
              preset.c_str(), cfg.nx, cfg.ny, cfg.lx, cfg.ly, (long long)n, cfg.dt, nsteps,
              prec.c_str());
  if (!gpu) {
    std::printf("GPU run skipped (-g); this build has no CPU path\n");
    return 0;
  }
  FILE* out = nullptr;
  if (!file.empty()) {
    out = std::fopen(file.c_str(), "w");
    if (!out) {
      std::fprintf(stderr, "Error: cannot open %s\n", file.c_str());
      return 1;
    }
    std::fprintf(out, "step,time,field_energy,kinetic_energy,total_energy,px,py,mode1,"
                      "particles,collisions\n");
  }
  const double t_init = now_s();
  pic_ctx* ctx = nullptr;
  if (!check(pic_create(&cfg, &ctx), "pic_create")) return 1;
  if (!check(pic_set_particles(ctx, 0, x.data(), y.data(), vx.data(), vy.data(), n),
             "pic_set_particles")) {
    pic_destroy(ctx);
    return 1;
  }
  if (autotune && !check(pic_autotune(ctx, 8, nullptr), "pic_autotune")) {
    pic_destroy(ctx);
    return 1;
  }
  if (!check(pic_sync(ctx), "pic_sync")) {
    pic_destroy(ctx);
    return 1;
  }
  const double init_s = now_s() - t_init;
  // step in chunks so the diagnostics ring (65536 rows) is drained in time
  const int chunk = 4096;
  std::vector<double> diag((size_t)chunk * PIC_DIAG_COLS);
  double gpu_s = 0.0;
  long long done = 0;
  bool ok = true;
  while (done < nsteps && ok) {
    const int k = (int)std::min<long long>(chunk, nsteps - done);
    const double t0 = now_s();
    ok = check(pic_step(ctx, k), "pic_step") && check(pic_sync(ctx), "pic_sync");
    gpu_s += now_s() - t0;
    if (ok && out) {
      ok = check(pic_get_diagnostics(ctx, 0, done, k, diag.data()), "pic_get_diagnostics");
      for (int s = 0; ok && s < k; ++s) {
        const double* d = &diag[(size_t)s * PIC_DIAG_COLS];
        std::fprintf(out, "%lld,%.9g,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g,%.0f,%.0f\n",
                     done + s, (done + s) * cfg.dt, d[0], d[1], d[0] + d[1], d[2], d[3], d[4],
                     d[5], d[6]);
      }
    }
    done += k;
  }
  if (out) std::fclose(out);
  if (ok && !final_file.empty()) {
    int64_t nf = 0;
    ok = check(pic_num_particles(ctx, 0, &nf), "pic_num_particles");
    std::vector<double> fx(nf), fy(nf), fvx(nf), fvy(nf);
    ok = ok && check(pic_get_particles(ctx, 0, fx.data(), fy.data(), fvx.data(), fvy.data(), nf, 1),
                     "pic_get_particles");
    FILE* fo = ok ? std::fopen(final_file.c_str(), "w") : nullptr;
    if (ok && !fo) {
      std::fprintf(stderr, "Error: cannot open %s\n", final_file.c_str());
      ok = false;
    }
    if (fo) {
      std::fprintf(fo, "x,y,vx,vy\n");
      for (int64_t i = 0; i < nf; ++i)
        std::fprintf(fo, "%.17g,%.17g,%.17g,%.17g\n", fx[i], fy[i], fvx[i], fvy[i]);
      std::fclose(fo);
    }
  }
  int tile = 0, halo = 0, splits = 0, se = 0;
  pic_get_tuning(ctx, &tile, &halo, &splits, &se);
  pic_destroy(ctx);
  if (!ok) return 1;
  std::printf("\nTuning: tile %d, halo %d, splits %d, sort every %d\n", tile, halo, splits, se);
  std::printf("Problem size: %lld particles, %d x %d grid\n", (long long)n, cfg.nx, cfg.ny);
  std::printf("Initialisation took %f seconds\n", init_s);
  std::printf("GPU took %f seconds (%.4g particle-steps/s)\n\n", gpu_s,
              gpu_s > 0 ? (double)n * nsteps / gpu_s : 0.0);
  return 0;
}
