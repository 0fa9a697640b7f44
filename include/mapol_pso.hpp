// mapol_pso.hpp -- the reference's particle-swarm loss minimiser (pso::run,
// src/pso/run.h:27-168; pso::Particula, src/pso/Particula.h:20-190; Solucao,
// Parametros, src/pso/Solucao.h, Parametros.h; the control vector built in
// src/main.cu:232-264) over the batched engine, through the C++ mirror of
// runpf.h (mapol_runpf.hpp).  A caller of the reference switches by include:
//
//   reference                                   here
//   pso::Particula::Estrutura                   mapol::Estrutura
//   pso::Particula / Solucao / Parametros       mapol::pso::Particula / ...
//   pso::run(&melhores, w_max, w_min, c, s,     mapol::pso::run(...)  same
//            nParticulas, nIteracoes,           arguments; execucao 'P' (one
//            tamParticula, &estrutura)          batched solve of the swarm per
//                                               iteration, hybrid_runpf) or
//                                               'S'/'O' (mkl_runpf per particle)
//
// The draws are the reference's: rand()/RAND_MAX and rand() % range
// (src/util/random.h), seeded by the caller (main.cu seeds with time(0)).
// run() takes an optional evaluator in place of the power flow (the role of
// the reference's test::fitness hook, run.h:11-14).  The solution history is
// returned without the reference's console output.
#pragma once

#include <chrono>
#include <climits>
#include <cstdlib>
#include <functional>
#include <sstream>
#include <string>
#include <vector>

#include "mapol_runpf.hpp"

namespace mapol {
namespace pso {

// src/util/random.h
inline double real_rand(double lo = 0.0, double hi = 1.0) {
  return (double)std::rand() / RAND_MAX * (hi - lo) + lo;
}
inline int rand_range(int lo, int hi) { return std::rand() % (hi - lo) + lo; }

// src/pso/Parametros.h (defaults of main.cu:26-31)
struct Parametros {
  double w_max = 0.9, w_min = 0.4, c = 2.0, s = 2.0;
  unsigned nParticulas = 1, nIteracoes = 200, tamParticula = 0;
};

// src/pso/Solucao.h: a global best as it was found
class Solucao {
 public:
  Solucao() = default;
  Solucao(const std::vector<double> &valores, double fitness, double tempo,
          unsigned iteracao)
      : valores_(valores), fitness_(fitness), tempo_(tempo), iteracao_(iteracao) {}
  double operator[](unsigned i) const { return valores_[i]; }
  const std::vector<double> &valores() const { return valores_; }
  double fitness() const { return fitness_; }
  double tempo() const { return tempo_; }
  unsigned iteracao() const { return iteracao_; }
  operator std::string() const {
    std::ostringstream ss;
    ss << "Solucao:" << fitness_ << ';' << tempo_ << ';' << iteracao_ << ';';
    for (double v : valores_) ss << v << ';';
    return ss.str();
  }

 private:
  std::vector<double> valores_;
  double fitness_ = 0, tempo_ = 0;
  unsigned iteracao_ = 0;
};

// src/pso/Particula.h:20-190: position X, velocity V, personal best
class Particula {
 public:
  std::vector<double> X, V, bests;
  std::vector<double> *bests_global = nullptr;
  std::vector<Estrutura> *estrutura = nullptr;

  size_t size() const { return X.size(); }
  double fitness() const { return fitness_; }
  double fitMelhor() const { return fitMelhor_; }
  double operator[](unsigned i) const { return X[i]; }

  // Particula.h:88-131: velocity U(0,1); position at the defaults, or drawn:
  // AVR uniform in [min, max], OLTC and SHC on their discrete steps
  void inicializar(unsigned n, std::vector<double> *global,
                   std::vector<Estrutura> *est, bool defaults = false) {
    X.assign(n, 0.0);
    V.assign(n, 0.0);
    bests.assign(n, 0.0);
    bests_global = global;
    estrutura = est;
    for (unsigned i = 0; i < n; i++) {
      V[i] = real_rand();
      const Estrutura &e = est->at(i);
      double x;
      if (defaults) {
        x = e.defaultValue;
      } else if (e.tipo == Estrutura::AVR) {
        x = real_rand(e.min, e.max);
      } else if (e.tipo == Estrutura::OLTC) {
        x = rand_range(0, (int)e.qtd) * (e.max - e.min) / e.qtd + e.min;
      } else if (e.tipo == Estrutura::SHC) {
        x = rand_range(0, (int)e.qtd) * e.max / e.qtd;
      } else {
        x = real_rand();
      }
      X[i] = bests[i] = x;
    }
    fitMelhor_ = DBL_MAX;
  }

  // Particula.h:133-145: keeps the personal best
  void mudarFitness(double v) {
    fitness_ = v;
    if (fitness_ < fitMelhor_) {
      fitMelhor_ = fitness_;
      bests = X;
    }
  }

  // Particula.h:147-187: velocity from inertia, personal and global pulls;
  // a position past a bound is reflected back by a random fraction of the
  // excess (repeatedly), then OLTC and SHC snap to their nearest step
  void mover(double w, double c, double s) {
    for (size_t i = 0; i < X.size(); i++) {
      const Estrutura &e = estrutura->at(i);
      const double r1 = real_rand(), r2 = real_rand();
      V[i] = w * V[i] + r1 * c * (bests[i] - X[i]) +
             r2 * s * (bests_global->at(i) - X[i]);
      X[i] += V[i];
      while (X[i] > e.max || X[i] < e.min) {
        if (X[i] > e.max) X[i] = e.max - real_rand() * (X[i] - e.max);
        else X[i] = e.min + real_rand() * (e.min - X[i]);
      }
      if (e.tipo == Estrutura::OLTC) {
        const int k = (int)((X[i] - e.min) * e.qtd / (e.max - e.min) + 0.5);
        X[i] = k * (e.max - e.min) / e.qtd + e.min;
      } else if (e.tipo == Estrutura::SHC) {
        const int k = (int)(X[i] * e.qtd / e.max + 0.5);
        X[i] = k * e.max / e.qtd;
      }
    }
  }

 private:
  double fitness_ = DBL_MAX, fitMelhor_ = DBL_MAX;
};

// main.cu:232-264: one control per PV bus (AVR set-point in [0.9, 1.1],
// default its VG), per bus with a shunt susceptance (SHC in [0, 0.18] pu,
// 3 steps, default its Bsh) and per branch with a tap (OLTC in [0.9, 1.1],
// 20 steps, default its tap); records indiceEstrutura on the topology
inline std::vector<Estrutura> estruturaPadrao(Topology &t) {
  std::vector<Estrutura> est;
  auto add = [&](Estrutura::Tipo tipo, double lo, double hi, double def,
                 unsigned qtd, unsigned id) {
    Estrutura e;
    e.tipo = tipo;
    e.min = lo;
    e.max = hi;
    e.defaultValue = def;
    e.qtd = qtd;
    e.id = id;
    est.push_back(e);
  };
  for (size_t i = 0; i < t.buses.size(); i++) {
    Bus &b = t.buses[i];
    if (b.type == Bus::PV) {
      b.indiceEstrutura = (int)est.size();
      add(Estrutura::AVR, 0.9, 1.1, b.VG, 0, (unsigned)i);
    }
    if (b.Bsh > 0) {
      b.indiceEstrutura = (int)est.size();
      add(Estrutura::SHC, 0.0, 0.18, b.Bsh, 3, (unsigned)i);
    }
  }
  for (size_t i = 0; i < t.branches.size(); i++) {
    Branch &br = t.branches[i];
    if (br.tap > 0) {
      br.indiceEstrutura = (int)est.size();
      add(Estrutura::OLTC, 0.9, 1.1, br.tap, 20, (unsigned)i);
    }
  }
  return est;
}

// run.h:27-53: the swarm's best personal best, when it beats the global one,
// becomes the global best and is recorded
inline void atualizaMelhor(std::vector<Solucao> *melhores,
                           std::vector<double> *bests_global, double *fitMg,
                           std::vector<Particula> &enxame, unsigned it,
                           double tempo) {
  int id = -1;
  for (size_t i = 0; i < enxame.size(); i++)
    if (enxame[i].fitMelhor() < *fitMg) {
      id = (int)i;
      *fitMg = enxame[i].fitMelhor();
    }
  if (id >= 0) {
    *bests_global = enxame[id].bests;
    melhores->push_back(Solucao(*bests_global, *fitMg, tempo, it));
  }
}

// the swarm's fitness: by default the power flow through the mirror
using Avaliador = std::function<void(std::vector<Estrutura> &, std::vector<Particula> &)>;

// run.h:58-168: particle 0 at the defaults, the others drawn; evaluation;
// then per iteration the inertia w_max - (w_max - w_min) it / nIteracoes, a
// move of every particle and an evaluation of the swarm.  execucao 'P'
// evaluates the swarm in one batched solve (hybrid_runpf: hybrid_init must
// have been called), 'S'/'O' one particle per call (mkl_runpf after
// mkl_init).
inline void run(std::vector<Solucao> *melhores, double w_max, double w_min,
                double c, double s, unsigned nParticulas, unsigned nIteracoes,
                unsigned tamParticula, std::vector<Estrutura> *estrutura,
                char execucao = 'P', Avaliador avaliar = nullptr) {
  const auto t0 = std::chrono::steady_clock::now();
  auto tempo = [&]() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  };
  if (!avaliar) {
    if (execucao == 'P') {
      avaliar = [](std::vector<Estrutura> &est, std::vector<Particula> &enx) {
        hybrid_runpf(est, enx);
      };
    } else if (execucao == 'S' || execucao == 'O') {
      avaliar = [](std::vector<Estrutura> &est, std::vector<Particula> &enx) {
        for (Particula &p : enx) p.mudarFitness(mkl_runpf(est, p));
      };
    } else {
      throw Error(std::string("pso::run: unknown execution type ") + execucao);
    }
  }
  std::vector<double> bests_global;
  std::vector<Particula> enxame(nParticulas);
  enxame[0].inicializar(tamParticula, &bests_global, estrutura, true);
  for (unsigned i = 1; i < nParticulas; i++)
    enxame[i].inicializar(tamParticula, &bests_global, estrutura);
  avaliar(*estrutura, enxame);
  double fitMg = UINT_MAX / 1e6;      // the reference's starting bar (run.h:118)
  bests_global = enxame[0].bests;
  atualizaMelhor(melhores, &bests_global, &fitMg, enxame, 0, tempo());
  for (unsigned it = 1; it <= nIteracoes; it++) {
    const double w = w_max - (w_max - w_min) / nIteracoes * it;
    for (Particula &p : enxame) p.mover(w, c, s);
    avaliar(*estrutura, enxame);
    atualizaMelhor(melhores, &bests_global, &fitMg, enxame, it, tempo());
  }
}

inline void run(std::vector<Solucao> *melhores, const Parametros &p,
                std::vector<Estrutura> *estrutura, char execucao = 'P',
                Avaliador avaliar = nullptr) {
  run(melhores, p.w_max, p.w_min, p.c, p.s, p.nParticulas, p.nIteracoes,
      p.tamParticula, estrutura, execucao, avaliar);
}

}  // namespace pso
}  // namespace mapol
