// Drives the engine through the C++ mirror of the reference's power-flow
// module (include/mapol_runpf.hpp) the way pso::run drives runpf.h: a
// 3-bus case, a swarm of particles whose AVR/OLTC controls vary, fitness =
// loss or DBL_MAX.  Prints one line per particle.
#include <cstdio>

#include "mapol_runpf.hpp"

// with a MATPOWER case file as argument: the case read through
// mapol::loadMatpowerFormat, two particles without controls (base-case loss)
static int run_case_file(const char *path) {
  try {
    mapol::Topology t = mapol::loadMatpowerFormat(path);
    std::vector<mapol::Estrutura> est;
    std::vector<mapol::Particula> swarm(2);
    mapol::hybrid_init(t, 2, 1, est, 0);
    mapol::hybrid_runpf(est, swarm);
    mapol::hybrid_free();
    std::printf("buses %zu branches %zu gens %zu\n", t.buses.size(),
                t.branches.size(), t.gens.size());
    for (int k = 0; k < 2; k++) std::printf("particle %d loss %.12e\n", k, swarm[k].fitness);
  } catch (const mapol::Error &e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}

int main(int argc, char **argv) {
  if (argc > 1) return run_case_file(argv[1]);
  mapol::Topology t;
  t.baseMVA = 100;
  t.buses.resize(3);
  t.buses[0].type = mapol::Bus::SLACK; t.buses[0].V = 1.02;
  t.buses[1].type = mapol::Bus::PV;    t.buses[1].V = 1.0; t.buses[1].VG = 1.01;
  t.buses[1].P = 0.4;
  t.buses[2].type = mapol::Bus::PQ;    t.buses[2].P = -0.9; t.buses[2].Q = -0.3;
  mapol::Branch b;
  b.R = 0.02; b.X = 0.08; b.B = 0.02; b.tap = 0;
  b.from = 0; b.to = 1; t.branches.push_back(b);
  b.from = 1; b.to = 2; t.branches.push_back(b);
  b.from = 0; b.to = 2; b.tap = 0.98; t.branches.push_back(b);
  mapol::Generator g;
  g.busID = 0; t.gens.push_back(g);
  g.busID = 1; g.P = 0.4; t.gens.push_back(g);
  std::vector<mapol::Estrutura> est(2);
  est[0].tipo = mapol::Estrutura::AVR;  est[0].id = 1;
  est[1].tipo = mapol::Estrutura::OLTC; est[1].id = 2;
  std::vector<mapol::Particula> swarm(4);
  for (int k = 0; k < 4; k++) swarm[k].X = {0.98 + 0.02 * k, 0.95 + 0.03 * k};
  try {
    mapol::hybrid_init(t, 4, 1, est, 0);
    mapol::hybrid_runpf(est, swarm);
    mapol::hybrid_free();
  } catch (const mapol::Error &e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  for (int k = 0; k < 4; k++)
    std::printf("particle %d AVR %.3f OLTC %.3f loss %.12e\n", k, swarm[k].X[0],
                swarm[k].X[1], swarm[k].fitness);
  return 0;
}
