// mapol_runpf.hpp -- C++ host interface over the batchpf C-ABI that mirrors
// the reference's power-flow module (src/powerflow/runpf.h) so that its
// caller (pso::run, src/pso/run.h:58-168) can switch to the MI355X engine by
// changing an include:
//
//   reference (runpf.h)                         here
//   hybrid_init(Topology&, nTest, nThreads,     mapol::hybrid_init(...)  same
//               vector<Estrutura>, algPF)       arguments and meaning
//   hybrid_runpf(vector<Estrutura>&,            mapol::hybrid_runpf(...) sets
//                vector<Particula>&)            every particle's fitness to its
//                                               loss, or DBL_MAX when the power
//                                               flow does not converge
//                                               (runpf.h:781-790)
//   hybrid_free()                               mapol::hybrid_free()
//   mkl_init / mkl_runpf / mkl_clean            mapol::mkl_* (same engine, one
//                                               instance per call)
//   loadMatpowerFormat(path)                    mapol::loadMatpowerFormat(path)
//   (src/files/loadMatpowerFormat.h:14)
//
// Differences kept on purpose: errors throw mapol::Error instead of exit()
// (helper_cuda.h:967-977); the state lives in one handle instead of module
// globals, and nThreads (a CPU thread count the reference also used as a GPU
// block size, runpf.h:804) is accepted and ignored.
#pragma once

#include <cfloat>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "batchpf.h"

namespace mapol {

struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// src/powersystem/Bus.h:6-63 (P, Q are net injections in pu, O in radians)
struct Bus {
  enum Type { NONE, PQ, PV, SLACK, ISOLATED };
  unsigned id = 0;
  Type type = NONE;
  double V = 1, VG = 1, Vmax = 1, Vmin = 0, baseKV = 1, O = 0, P = 0, Q = 0,
         Gsh = 0, Bsh = 0;
  int indicePVPQ = -1, indiceEstrutura = -1;
};

// src/powersystem/Branch.h:10-62
struct Branch {
  unsigned from = 0, to = 0;
  bool inservice = true;
  double R = 0, X = 0, B = 0, tap = 1, shift = 0;
  int indiceEstrutura = -1;
};

// src/powersystem/Generator.h:9-40
struct Generator {
  unsigned busID = 0;
  double P = 0, Q = 0, Pmax = 0, Pmin = 0, Qmax = 0, Qmin = 0;
};

// src/powersystem/Topology.h:18-127 (lists of PV/PQ ids are derived)
struct Topology {
  double baseMVA = 100;
  std::vector<Bus> buses;
  std::vector<Branch> branches;
  std::vector<Generator> gens;
};

// src/pso/Particula.h:26-60
struct Estrutura {
  enum Tipo { NULO, AVR, OLTC, SHC };
  Tipo tipo = NULO;
  double min = -DBL_MAX, max = DBL_MAX, defaultValue = 0;
  unsigned qtd = 0, id = 0;
};

// the part of pso::Particula the power flow touches (Particula.h:73-145)
struct Particula {
  std::vector<double> X;
  double fitness = DBL_MAX;
  double operator[](unsigned i) const { return X[i]; }
  void mudarFitness(double v) { fitness = v; }
};

// src/files/loadMatpowerFormat.h:14-209 through bpf_load_matpower: bus P/Q
// are net injections (sum of generation minus load, pu), O in radians, VG
// from the last generator of the bus, ids in file order.  Generator limits
// are kept in pu (the reference's Generator keeps MW; its power flow never
// reads them).  Throws mapol::Error where the reference calls exit().
inline Topology loadMatpowerFormat(const std::string &path) {
  bpf_case c;
  const double *rate = nullptr;
  bpf_case_data *owner = nullptr;
  bpf_status s = bpf_load_matpower(path.c_str(), 0, &c, &rate, &owner);
  if (s != BPF_OK) {
    std::string msg = owner ? bpf_case_error(owner) : "bpf_load_matpower";
    bpf_free_case(owner);
    throw Error("loadMatpowerFormat: " + msg);
  }
  Topology t;
  t.baseMVA = c.baseMVA;
  t.buses.resize(c.nbus);
  for (int i = 0; i < c.nbus; i++) {
    Bus &b = t.buses[i];
    b.id = (unsigned)i;
    b.type = (Bus::Type)c.bus_type[i];
    b.V = c.Vm[i];
    b.O = c.Va[i];
    b.P = -c.Pd[i];
    b.Q = -c.Qd[i];
    b.Gsh = c.Gs[i];
    b.Bsh = c.Bs[i];
  }
  t.gens.resize(c.ngen);
  for (int k = 0; k < c.ngen; k++) {
    Generator &g = t.gens[k];
    g.busID = (unsigned)c.gen_bus[k];
    g.P = c.Pg[k];
    g.Q = c.Qg[k];
    g.Qmax = c.Qmax[k];
    g.Qmin = c.Qmin[k];
    t.buses[g.busID].P += c.Pg[k];
    t.buses[g.busID].Q += c.Qg[k];
    t.buses[g.busID].VG = c.VG[k];
  }
  t.branches.resize(c.nbranch);
  for (int k = 0; k < c.nbranch; k++) {
    Branch &b = t.branches[k];
    b.from = (unsigned)c.br_from[k];
    b.to = (unsigned)c.br_to[k];
    b.R = c.R[k];
    b.X = c.X[k];
    b.B = c.B[k];
    b.tap = c.tap[k];
    b.shift = c.shift[k];
    b.inservice = c.br_status[k] != 0;
  }
  bpf_free_case(owner);
  return t;
}

namespace detail {
struct State {
  bpf_engine *eng = nullptr;
  int nTest = 0, nbus = 0;
  std::vector<bpf_ctl> ctl;
};
inline State &state(bool cpu) {
  static State g, c;
  return cpu ? c : g;
}
inline void check(bpf_engine *e, bpf_status s, const char *what) {
  if (s != BPF_OK)
    throw Error(std::string(what) + ": " + (e ? bpf_last_error(e) : "no engine"));
}
inline void init(State &S, Topology &t, int nTest,
                 const std::vector<Estrutura> &est, int algPF) {
  const int n = (int)t.buses.size(), m = (int)t.branches.size(),
            g = (int)t.gens.size();
  std::vector<int32_t> type(n), from(m), to(m), gbus(g);
  std::vector<double> Pd(n), Qd(n), Gs(n), Bs(n), Vm(n), Va(n), R(m), X(m),
      B(m), tap(m), shift(m), zero(g, 0.0), Qmax(g), Qmin(g), VG(g);
  std::vector<uint8_t> st(m), gst(g, 1);
  for (int i = 0; i < n; i++) {
    const Bus &b = t.buses[i];
    type[i] = (int32_t)b.type;
    // the reference stores net injections; the engine takes loads and
    // generation apart, so the net injection goes in as a negative load
    Pd[i] = -b.P;
    Qd[i] = -b.Q;
    Gs[i] = b.Gsh;
    Bs[i] = b.Bsh;
    Vm[i] = b.V;
    Va[i] = b.O;
  }
  for (int k = 0; k < m; k++) {
    const Branch &b = t.branches[k];
    from[k] = (int32_t)b.from;
    to[k] = (int32_t)b.to;
    R[k] = b.R; X[k] = b.X; B[k] = b.B; tap[k] = b.tap; shift[k] = b.shift;
    st[k] = b.inservice ? 1 : 0;
  }
  for (int k = 0; k < g; k++) {
    gbus[k] = (int32_t)t.gens[k].busID;
    Qmax[k] = t.gens[k].Qmax;
    Qmin[k] = t.gens[k].Qmin;
    VG[k] = t.buses[gbus[k]].VG;
  }
  bpf_case c{n, m, g, t.baseMVA, type.data(), Pd.data(), Qd.data(), Gs.data(),
             Bs.data(), Vm.data(), Va.data(), from.data(), to.data(), R.data(),
             X.data(), B.data(), tap.data(), shift.data(), st.data(),
             gbus.data(), zero.data(), zero.data(), Qmax.data(), Qmin.data(),
             VG.data(), gst.data()};
  bpf_options o;
  bpf_default_options(&o);
  o.alg = algPF;
  if (S.eng) { bpf_destroy(S.eng); S.eng = nullptr; }
  bpf_status s = bpf_create(&c, &o, &S.eng);
  if (s != BPF_OK) {
    std::string msg = S.eng ? bpf_last_error(S.eng) : "bpf_create";
    bpf_destroy(S.eng);
    S.eng = nullptr;
    throw Error("hybrid_init: " + msg);
  }
  S.nTest = nTest;
  S.nbus = n;
  S.ctl.clear();
  for (const Estrutura &e : est) {
    int kind = e.tipo == Estrutura::AVR ? BPF_CTL_AVR
               : e.tipo == Estrutura::OLTC ? BPF_CTL_OLTC
               : e.tipo == Estrutura::SHC ? BPF_CTL_SHC : 0;
    if (!kind) throw Error("hybrid_init: control of type NULO");
    S.ctl.push_back(bpf_ctl{kind, (int32_t)e.id});
  }
}
// losses of the swarm's control vectors (any particle type with operator[]
// over its controls), DBL_MAX where the power flow does not converge
template <class P>
inline std::vector<double> losses(State &S, const std::vector<P> &swarm) {
  if (!S.eng) throw Error("runpf before init");
  const int T = (int)swarm.size(), nc = (int)S.ctl.size();
  std::vector<double> vals((size_t)T * nc);
  for (int t = 0; t < T; t++)
    for (int c = 0; c < nc; c++) vals[(size_t)t * nc + c] = swarm[t][c];
  check(S.eng, bpf_set_instances(S.eng, T, nullptr, nullptr, nc,
                                 nc ? S.ctl.data() : nullptr,
                                 nc ? vals.data() : nullptr),
        "bpf_set_instances");
  check(S.eng, bpf_solve(S.eng), "bpf_solve");
  std::vector<double> loss(T);
  check(S.eng, bpf_get_results(S.eng, nullptr, nullptr, nullptr, nullptr,
                               loss.data(), nullptr, nullptr),
        "bpf_get_results");
  return loss;
}
}  // namespace detail

// runpf.h:795 -- nThreads is accepted for signature parity and ignored
inline void hybrid_init(Topology &topology, int nTest, int /*nThreads*/,
                        std::vector<Estrutura> estrutura, int algPF) {
  detail::init(detail::state(false), topology, nTest, estrutura, algPF);
}

// runpf.h:606 -- all particles in one batched solve; each particle's
// mudarFitness receives its loss (runpf.h:782-790).  P: Particula or any
// particle type with operator[] and mudarFitness (the swarm of mapol_pso.hpp)
template <class P>
inline void hybrid_runpf(std::vector<Estrutura> & /*estrutura*/,
                         std::vector<P> &enxame) {
  const std::vector<double> loss = detail::losses(detail::state(false), enxame);
  for (size_t t = 0; t < enxame.size(); t++) enxame[t].mudarFitness(loss[t]);
}

// runpf.h:984
inline void hybrid_free() {
  auto &S = detail::state(false);
  bpf_destroy(S.eng);
  S.eng = nullptr;
}

// runpf.h:359, 259, 452 -- one particle per call
inline void mkl_init(Topology &topology, int nTest,
                     std::vector<Estrutura> estrutura, int algPF) {
  detail::init(detail::state(true), topology, nTest, estrutura, algPF);
}
// returns the loss; the caller hands it to mudarFitness (main.cu:300-306)
template <class P>
inline double mkl_runpf(std::vector<Estrutura> & /*estrutura*/,
                        P &particula) {
  return detail::losses(detail::state(true), std::vector<P>(1, particula))[0];
}
inline void mkl_clean() {
  auto &S = detail::state(true);
  bpf_destroy(S.eng);
  S.eng = nullptr;
}

}  // namespace mapol
