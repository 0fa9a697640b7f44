// Drives mapol::pso::run (the mirror of the reference's pso::run).
//
//   pso_demo                      the swarm mechanics on a known fitness (no
//                                 GPU): a sum of squares around a target on
//                                 the controls' grids; prints every recorded
//                                 global best and the checks of each move
//   pso_demo case.m seed P I      loss minimisation of a MATPOWER case with
//                                 the reference's control vector (AVR, SHC,
//                                 OLTC), P particles, I iterations, one
//                                 batched solve per iteration; prints the
//                                 recorded global bests and the final controls
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "mapol_pso.hpp"

using mapol::Estrutura;
using mapol::pso::Particula;

static bool on_grid(const Estrutura &e, double x) {
  if (e.tipo == Estrutura::OLTC) {
    const double k = (x - e.min) * e.qtd / (e.max - e.min);
    return std::fabs(k - std::round(k)) < 1e-9;
  }
  if (e.tipo == Estrutura::SHC) {
    const double k = x * e.qtd / e.max;
    return std::fabs(k - std::round(k)) < 1e-9;
  }
  return true;
}

static int mechanics() {
  std::vector<Estrutura> est(5);
  const Estrutura::Tipo tipos[5] = {Estrutura::AVR, Estrutura::AVR, Estrutura::OLTC,
                                    Estrutura::OLTC, Estrutura::SHC};
  const double lo[5] = {0.9, 0.9, 0.9, 0.9, 0.0}, hi[5] = {1.1, 1.1, 1.1, 1.1, 0.18};
  const unsigned qtd[5] = {0, 0, 20, 20, 3};
  for (int i = 0; i < 5; i++) {
    est[i].tipo = tipos[i];
    est[i].min = lo[i];
    est[i].max = hi[i];
    est[i].qtd = qtd[i];
    est[i].defaultValue = tipos[i] == Estrutura::SHC ? 0.0 : 1.0;
    est[i].id = i;
  }
  const double target[5] = {1.03, 0.97, 1.05, 0.93, 0.12};  // on the grids
  long evaluations = 0, out_of_bounds = 0, off_grid = 0, stale_best = 0;
  auto fitness = [&](const std::vector<double> &x) {
    double f = 0;
    for (int i = 0; i < 5; i++) f += (x[i] - target[i]) * (x[i] - target[i]);
    return f;
  };
  mapol::pso::Avaliador avaliar = [&](std::vector<Estrutura> &e,
                                     std::vector<Particula> &enx) {
    for (Particula &p : enx) {
      for (size_t i = 0; i < p.size(); i++) {
        // the step formula of OLTC/SHC may land a rounding above max
        if (p.X[i] < e[i].min - 1e-12 || p.X[i] > e[i].max + 1e-12) out_of_bounds++;
        if (!on_grid(e[i], p.X[i])) off_grid++;
      }
      const double before = p.fitMelhor();
      p.mudarFitness(fitness(p.X));
      // the personal best is the best position seen
      if (p.fitMelhor() > before || std::fabs(fitness(p.bests) - p.fitMelhor()) > 0) stale_best++;
      evaluations++;
    }
  };
  std::srand(12345);
  std::vector<mapol::pso::Solucao> melhores;
  mapol::pso::Parametros par;
  par.nParticulas = 24;
  par.nIteracoes = 60;
  par.tamParticula = 5;
  mapol::pso::run(&melhores, par, &est, 'P', avaliar);
  for (const auto &s : melhores)
    std::printf("best it %u fitness %.17g\n", s.iteracao(), s.fitness());
  std::printf("evaluations %ld out_of_bounds %ld off_grid %ld stale_best %ld\n",
              evaluations, out_of_bounds, off_grid, stale_best);
  const auto &b = melhores.back();
  std::printf("final");
  for (int i = 0; i < 5; i++) std::printf(" %.17g", b[i]);
  std::printf("\nsolution %s\n", std::string(b).c_str());
  return 0;
}

static int case_file(const char *path, unsigned seed, unsigned np, unsigned nit) {
  try {
    mapol::Topology t = mapol::loadMatpowerFormat(path);
    std::vector<Estrutura> est = mapol::pso::estruturaPadrao(t);
    std::srand(seed);
    mapol::hybrid_init(t, (int)np, 1, est, 0);
    // the defaults' loss (particle 0 of the swarm starts there)
    std::vector<Particula> base(1);
    std::vector<double> none;
    base[0].inicializar((unsigned)est.size(), &none, &est, true);
    mapol::hybrid_runpf(est, base);
    std::printf("default loss %.17g\n", base[0].fitness());
    std::vector<mapol::pso::Solucao> melhores;
    mapol::pso::run(&melhores, 0.9, 0.4, 2.0, 2.0, np, nit,
                    (unsigned)est.size(), &est);
    // a probe away from the defaults: every control at its range's midpoint
    // step (AVR 1.0, OLTC 1.0, SHC max/3)
    std::vector<Particula> probe(1);
    probe[0].inicializar((unsigned)est.size(), &none, &est, true);
    for (size_t i = 0; i < est.size(); i++)
      probe[0].X[i] = est[i].tipo == Estrutura::SHC ? est[i].max / 3 : 1.0;
    mapol::hybrid_runpf(est, probe);
    mapol::hybrid_free();
    for (const auto &s : melhores)
      std::printf("best it %u fitness %.17g\n", s.iteracao(), s.fitness());
    const auto &b = melhores.back();
    for (size_t i = 0; i < est.size(); i++)
      std::printf("control %d %u %.17g %.17g\n", (int)est[i].tipo, est[i].id,
                  b[(unsigned)i], est[i].defaultValue);
    std::printf("probe loss %.17g\n", probe[0].fitness());
    for (size_t i = 0; i < est.size(); i++)
      std::printf("probe %d %u %.17g\n", (int)est[i].tipo, est[i].id, probe[0].X[i]);
  } catch (const mapol::Error &e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}

int main(int argc, char **argv) {
  if (argc > 1)
    return case_file(argv[1], argc > 2 ? (unsigned)atoi(argv[2]) : 1u,
                     argc > 3 ? (unsigned)atoi(argv[3]) : 16u,
                     argc > 4 ? (unsigned)atoi(argv[4]) : 10u);
  return mechanics();
}
