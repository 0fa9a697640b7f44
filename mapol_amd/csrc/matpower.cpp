// MATPOWER (version 2) case reader behind the C-ABI: bpf_load_matpower.
//
// Same conventions as the reference's loader (src/files/loadMatpowerFormat.h
// :14-209):
//  * bus rows  id type Pd Qd Gs Bs area Vm Va baseKV zone Vmax Vmin; loads and
//    shunts divided by baseMVA, Va converted to radians (:78-90);
//  * bus ids mapped to 0-based positions in file order (:91, :116-122,
//    :184-198);
//  * gen rows  bus Pg Qg Qmax Qmin Vg mBase status Pmax Pmin; every
//    generator's Pg/Qg is added to its bus whatever its status column says
//    (:125-128), the last generator of a bus sets its VG (:139-140);
//  * branch rows  f t r x b rateA rateB rateC ratio angle status (:167-183).
// Differences, all explicit: an unknown bus id is an error (the reference
// keeps the raw id and indexes out of range); Qmax = Inf / Qmin = -Inf become
// +-infinity (the reference stores +-DBL_MAX, which compares the same in
// every limit check); loads and generation stay separate (the reference
// stores the net injection Pg - Pd); BPF_MP_GEN_STATUS honours the generator
// status column (off by default, as in the reference).
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "batchpf.h"

struct bpf_case_data {
  std::vector<int32_t> bus_type, br_from, br_to, gen_bus;
  std::vector<double> Pd, Qd, Gs, Bs, Vm, Va, R, X, B, tap, shift, rate_a;
  std::vector<double> Pg, Qg, Qmax, Qmin, VG;
  std::vector<uint8_t> br_status, gen_status;
  std::string err;
};

namespace {

// rows of one "mpc.<name> = [ ... ];" block: '%' comments dropped, rows end
// at ';' or at a line end
bool read_block(const std::string &text, const char *name,
                std::vector<std::vector<double>> &rows, std::string &err) {
  rows.clear();
  const std::string key = std::string("mpc.") + name;
  size_t at = 0;
  for (;;) {
    at = text.find(key, at);
    if (at == std::string::npos) return true;  // block absent
    size_t q = at + key.size();
    while (q < text.size() && (text[q] == ' ' || text[q] == '\t')) q++;
    if (q < text.size() && text[q] == '=') { at = q; break; }
    at = q;
  }
  size_t open = text.find('[', at);
  if (open == std::string::npos) { err = std::string("mpc.") + name + ": no '['"; return false; }
  std::vector<double> row;
  size_t i = open + 1;
  std::string tok;
  auto flush_tok = [&]() -> bool {
    if (tok.empty()) return true;
    double v;
    if (tok == "Inf" || tok == "inf" || tok == "+Inf") v = INFINITY;
    else if (tok == "-Inf" || tok == "-inf") v = -INFINITY;
    else {
      char *end = nullptr;
      errno = 0;
      v = std::strtod(tok.c_str(), &end);
      if (end == tok.c_str() || *end != '\0') {
        err = std::string("mpc.") + name + ": bad number '" + tok + "'";
        return false;
      }
    }
    row.push_back(v);
    tok.clear();
    return true;
  };
  auto flush_row = [&]() {
    if (!row.empty()) rows.push_back(row);
    row.clear();
  };
  for (; i < text.size(); i++) {
    const char ch = text[i];
    if (ch == '%') {
      if (!flush_tok()) return false;
      while (i < text.size() && text[i] != '\n') i++;
      flush_row();
      continue;
    }
    if (ch == ']') {
      if (!flush_tok()) return false;
      flush_row();
      return true;
    }
    if (ch == ';' || ch == '\n' || ch == '\r') {
      if (!flush_tok()) return false;
      flush_row();
      continue;
    }
    if (ch == ' ' || ch == '\t' || ch == ',') {
      if (!flush_tok()) return false;
      continue;
    }
    tok.push_back(ch);
  }
  err = std::string("mpc.") + name + ": no closing ']'";
  return false;
}

bool scalar_after(const std::string &text, const char *key, std::string &out) {
  size_t at = text.find(key);
  if (at == std::string::npos) return false;
  size_t eq = text.find('=', at);
  size_t semi = text.find(';', eq);
  if (eq == std::string::npos || semi == std::string::npos) return false;
  out = text.substr(eq + 1, semi - eq - 1);
  size_t a = out.find_first_not_of(" \t'\"");
  size_t b = out.find_last_not_of(" \t'\"");
  out = (a == std::string::npos) ? "" : out.substr(a, b - a + 1);
  return true;
}

}  // namespace

extern "C" bpf_status bpf_load_matpower(const char *path, int32_t flags,
                                        bpf_case *out, const double **rate_a,
                                        bpf_case_data **owner) {
  if (!path || !out || !owner) return BPF_ERR_INVALID_ARG;
  *owner = nullptr;
  auto *D = new bpf_case_data();
  auto fail = [&](bpf_status st, const std::string &msg) {
    D->err = msg;
    *owner = D;  // keeps the message readable through bpf_case_error
    memset(out, 0, sizeof(*out));
    return st;
  };
  FILE *f = std::fopen(path, "rb");
  if (!f) return fail(BPF_ERR_IO, std::string("cannot open ") + path);
  std::string text;
  char buf[1 << 16];
  size_t got;
  while ((got = std::fread(buf, 1, sizeof buf, f)) > 0) text.append(buf, got);
  std::fclose(f);

  std::string ver, base;
  if (scalar_after(text, "mpc.version", ver) && ver != "2")
    return fail(BPF_ERR_UNSUPPORTED, "MATPOWER case input: old version");
  if (!scalar_after(text, "mpc.baseMVA", base))
    return fail(BPF_ERR_INVALID_ARG, "mpc.baseMVA missing");
  const double baseMVA = std::atof(base.c_str());
  if (!(baseMVA > 0)) return fail(BPF_ERR_INVALID_ARG, "mpc.baseMVA not positive");
  std::vector<std::vector<double>> bus, gen, br;
  std::string err;
  if (!read_block(text, "bus", bus, err) || !read_block(text, "gen", gen, err) ||
      !read_block(text, "branch", br, err))
    return fail(BPF_ERR_INVALID_ARG, err);
  if (bus.empty()) return fail(BPF_ERR_INVALID_ARG, "mpc.bus is empty");

  const int n = (int)bus.size(), m = (int)br.size(), g = (int)gen.size();
  std::unordered_map<long long, int> pos;
  for (int i = 0; i < n; i++) {
    if (bus[i].size() < 13) return fail(BPF_ERR_INVALID_ARG, "bus row with fewer than 13 columns");
    long long id = (long long)bus[i][0];
    if (!pos.emplace(id, i).second)
      return fail(BPF_ERR_INVALID_ARG, "duplicate bus id " + std::to_string(id));
  }
  auto where = [&](double id, int &k) {
    auto it = pos.find((long long)id);
    if (it == pos.end()) return false;
    k = it->second;
    return true;
  };
  D->bus_type.resize(n); D->Pd.resize(n); D->Qd.resize(n); D->Gs.resize(n);
  D->Bs.resize(n); D->Vm.resize(n); D->Va.resize(n);
  for (int i = 0; i < n; i++) {
    const auto &r = bus[i];
    D->bus_type[i] = (int32_t)r[1];
    D->Pd[i] = r[2] / baseMVA;
    D->Qd[i] = r[3] / baseMVA;
    D->Gs[i] = r[4] / baseMVA;
    D->Bs[i] = r[5] / baseMVA;
    D->Vm[i] = r[7];
    D->Va[i] = r[8] * M_PI / 180.0;
  }
  D->gen_bus.resize(g); D->Pg.resize(g); D->Qg.resize(g); D->Qmax.resize(g);
  D->Qmin.resize(g); D->VG.resize(g); D->gen_status.resize(g);
  for (int k = 0; k < g; k++) {
    const auto &r = gen[k];
    if (r.size() < 8) return fail(BPF_ERR_INVALID_ARG, "gen row with fewer than 8 columns");
    int b;
    if (!where(r[0], b))
      return fail(BPF_ERR_INVALID_ARG, "gen at unknown bus " + std::to_string((long long)r[0]));
    D->gen_bus[k] = b;
    D->Pg[k] = r[1] / baseMVA;
    D->Qg[k] = r[2] / baseMVA;
    D->Qmax[k] = r[3] / baseMVA;
    D->Qmin[k] = r[4] / baseMVA;
    D->VG[k] = r[5];
    D->gen_status[k] = (flags & BPF_MP_GEN_STATUS) ? (r[7] > 0) : 1;
  }
  D->br_from.resize(m); D->br_to.resize(m); D->R.resize(m); D->X.resize(m);
  D->B.resize(m); D->tap.resize(m); D->shift.resize(m); D->rate_a.resize(m);
  D->br_status.resize(m);
  for (int k = 0; k < m; k++) {
    const auto &r = br[k];
    if (r.size() < 11) return fail(BPF_ERR_INVALID_ARG, "branch row with fewer than 11 columns");
    int a, b;
    if (!where(r[0], a) || !where(r[1], b))
      return fail(BPF_ERR_INVALID_ARG, "branch " + std::to_string(k) + " at an unknown bus");
    D->br_from[k] = a;
    D->br_to[k] = b;
    D->R[k] = r[2];
    D->X[k] = r[3];
    D->B[k] = r[4];
    D->rate_a[k] = r[5] / baseMVA;
    D->tap[k] = r[8];
    D->shift[k] = r[9];
    D->br_status[k] = r[10] != 0.0;
  }
  bpf_case c;
  memset(&c, 0, sizeof c);
  c.nbus = n; c.nbranch = m; c.ngen = g; c.baseMVA = baseMVA;
  c.bus_type = D->bus_type.data();
  c.Pd = D->Pd.data(); c.Qd = D->Qd.data(); c.Gs = D->Gs.data(); c.Bs = D->Bs.data();
  c.Vm = D->Vm.data(); c.Va = D->Va.data();
  c.br_from = D->br_from.data(); c.br_to = D->br_to.data();
  c.R = D->R.data(); c.X = D->X.data(); c.B = D->B.data();
  c.tap = D->tap.data(); c.shift = D->shift.data();
  c.br_status = D->br_status.data();
  c.gen_bus = D->gen_bus.data(); c.Pg = D->Pg.data(); c.Qg = D->Qg.data();
  c.Qmax = D->Qmax.data(); c.Qmin = D->Qmin.data(); c.VG = D->VG.data();
  c.gen_status = D->gen_status.data();
  *out = c;
  if (rate_a) *rate_a = D->rate_a.data();
  *owner = D;
  return BPF_OK;
}

extern "C" const char *bpf_case_error(const bpf_case_data *owner) {
  return owner ? owner->err.c_str() : "";
}

extern "C" void bpf_free_case(bpf_case_data *owner) { delete owner; }
