// kdtree_nbody -- command-line driver of the kD-tree step, after
// Parallel/RustVersion/src/main.rs:8-31: generates circular_orbits(number)
// and runs simple_sim with dt = 1e-3 for `steps` steps, printing the elapsed
// seconds (initial conditions included, as there).  --dump writes the final
// state after the timed run, as seven arrays of N doubles (x, y, z, vx, vy,
// vz, m; native byte order), for checking a run against another
// implementation.
//
//   kdtree_nbody -n|--number N [-s|--steps S] [--theta T] [--leaf L]
//                [--mixed] [--ring] [--seed K] [--dump FILE]
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "kdtree.hpp"

static void usage() {
  std::fprintf(stderr,
               "usage: kdtree_nbody -n|--number N [-s|--steps S] [--theta T] [--leaf L] "
               "[--mixed] [--ring] [--seed K] [--dump FILE]\n");
  std::exit(2);
}

int main(int argc, char **argv) {
  long long number = -1, steps = 1;
  double theta = kdnbody::THETA;
  int leaf = (int)kdnbody::MAX_PARTS;
  int precision = KD_PREC_FP64;
  bool ring = false;
  unsigned long long seed = 0;
  const char *dump = nullptr;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto val = [&]() -> const char * {
      if (i + 1 >= argc) usage();
      return argv[++i];
    };
    if (a == "-n" || a == "--number") number = std::atoll(val());
    else if (a == "-s" || a == "--steps") steps = std::atoll(val());
    else if (a == "--theta") theta = std::atof(val());
    else if (a == "--leaf") leaf = std::atoi(val());
    else if (a == "--seed") seed = std::strtoull(val(), nullptr, 10);
    else if (a == "--mixed") precision = KD_PREC_MIXED;
    else if (a == "--ring") ring = true;
    else if (a == "--dump") dump = val();
    else usage();
  }
  if (number < 1 || steps < 0) usage();
  try {
    const auto start = std::chrono::steady_clock::now();
    std::vector<kdnbody::Particle> bodies =
        ring ? kdnbody::ring((size_t)number, seed) : kdnbody::circular_orbits((size_t)number, seed);
    kdnbody::simple_sim(bodies, 1e-3, steps, leaf, theta, precision);
    const double secs =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
    std::printf("%.9g\n", secs);
    if (dump) {
      FILE *f = std::fopen(dump, "wb");
      if (!f) {
        std::fprintf(stderr, "kdtree_nbody: cannot open %s\n", dump);
        return 1;
      }
      std::vector<double> col(bodies.size());
      bool ok = true;
      for (int q = 0; q < 7 && ok; ++q) {
        for (size_t i = 0; i < bodies.size(); ++i) {
          const kdnbody::Particle &b = bodies[i];
          col[i] = q < 3 ? b.p[q] : q < 6 ? b.v[q - 3] : b.m;
        }
        ok = std::fwrite(col.data(), sizeof(double), col.size(), f) == col.size();
      }
      if (std::fclose(f) != 0 || !ok) {
        std::fprintf(stderr, "kdtree_nbody: cannot write %s\n", dump);
        return 1;
      }
    }
  } catch (const kdnbody::Error &e) {
    std::fprintf(stderr, "kdtree_nbody: %s\n", e.what());
    return 1;
  }
  return 0;
}
