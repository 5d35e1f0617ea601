// Output and format layer: the reference's `kal` and `pro` files, its guess
// files, and a runner that takes an input file through its start blocks the
// way the reference program does.
//
//   kal   one row per solve: the requested scalars, `%.16e` for reals, `%i`
//         for integers, strings as they are, "NiN" for a quantity that is not
//         known; a header of `key<sep>name<sep>prop` names when the file is
//         new (Output::WriteOutput, src/output.cpp:511-557).
//   pro   the requested profiles with the coordinates of every interior site
//         (or every site with write_bounds): 1-D rows `%e` then `%.20g`
//         values, 2-D / 3-D rows `%e` coordinates then `%.20e` values, each
//         followed by a tab (LGrad1/LGrad2/LGrad3::PutProfiles,
//         src/LGrad1.cpp:228-249, src/LGrad2.cpp:185-205, src/LGrad3.cpp:81-102).
//   guess method, sizes and refinement, charge flag, field-block names, then
//         one `%e` value per line (Lattice::StoreGuess,
//         src/lattice.cpp:971-1019); read back by name (Lattice::ReadGuess and
//         Solve_scf::Guess, src/lattice.cpp:926-968, src/solve_scf.cpp:489-524).
//
// File naming follows the reference: the stem is the input file name up to
// its first '.', without directories (use_output_folder); `kal` files are
// <stem>.kal, profiles <stem>.pro for a single start block and
// <stem>_<start>.pro otherwise, guesses <stem>_<start>.outiv (or the
// guess_outputfile name with _<start> before its extension), all inside the
// output folder (./output/ unless `out_info : folder : path` says otherwise).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "scf_internal.h"

namespace nscf {
namespace {

// Lattice::WeightedSum with the ghosts removed (src/LGrad1.cpp:213-221,
// src/LGrad2.cpp:147-157, src/LGrad3.cpp:40-46): sums in site order.
double weighted_sum(const LatticeSpec &l, const std::vector<double> &X) {
  double s = 0;
  if (l.gradients == 3) {
    for (int i = 0; i < l.M; i++) s += X[i];
    return s;
  }
  if (l.geometry == GEOM_PLANAR) {
    for (int i = 0; i < l.M; i++) s += X[i];
    return l.gradients == 1 ? s / l.fjc : s / (l.fjc * l.fjc);
  }
  for (int i = 0; i < l.M; i++) s += X[i] * l.L[i];
  return s;
}

// zero the ghost layers (the reference's remove_bounds before its sums and
// before a profile is written)
void remove_ghosts(const LatticeSpec &l, std::vector<double> &X) {
  const int f = l.fjc;
  const int nx = l.MX + 2 * f, ny = l.gradients >= 2 ? l.MY + 2 * f : 1, nz = l.gradients == 3 ? l.MZ + 2 * f : 1;
  for (int x = 0; x < nx; x++)
    for (int y = 0; y < ny; y++)
      for (int z = 0; z < nz; z++) {
        const bool in = x >= f && x < l.MX + f && (l.gradients < 2 || (y >= f && y < l.MY + f)) &&
                        (l.gradients < 3 || (z >= f && z < l.MZ + f));
        if (!in) X[(size_t)x * l.JX + (size_t)y * l.JY + z * l.JZ] = 0.0;
      }
}

// ghost layers as the boundary conditions fill them (set_bounds: mirror and
// periodic ghosts copy their interior partner, a surface ghost keeps its own
// value), the state of a profile when the reference writes it
int ghost_map(int x, int n, int f, const std::vector<int> &lo, const std::vector<int> &hi) {
  if (x < f) return lo[x];
  if (x >= n + f) return hi[x - n - f];
  return x;
}

void set_ghosts(const LatticeSpec &l, std::vector<double> &X) {
  const int f = l.fjc;
  const int nx = l.MX + 2 * f, ny = l.gradients >= 2 ? l.MY + 2 * f : 1, nz = l.gradients == 3 ? l.MZ + 2 * f : 1;
  std::vector<int> lx = l.bx1f, hx = l.bxmf, ly = l.by1f, hy = l.bymf, lz = l.bz1f, hz = l.bzmf;
  if (f == 1) {
    lx = {l.BX1};
    hx = {l.BXM};
    ly = {l.BY1};
    hy = {l.BYM};
    lz = {l.BZ1};
    hz = {l.BZM};
  }
  const std::vector<double> Y = X;
  for (int x = 0; x < nx; x++)
    for (int y = 0; y < ny; y++)
      for (int z = 0; z < nz; z++) {
        const int sx = ghost_map(x, l.MX, f, lx, hx);
        const int sy = l.gradients >= 2 ? ghost_map(y, l.MY, f, ly, hy) : y;
        const int sz = l.gradients == 3 ? ghost_map(z, l.MZ, f, lz, hz) : z;
        X[(size_t)x * l.JX + (size_t)y * l.JY + z * l.JZ] = Y[(size_t)sx * l.JX + (size_t)sy * l.JY + sz * l.JZ];
      }
}

struct Value {
  int kind = 0;  // 0 unknown, 1 int, 2 real, 3 string
  int i = 0;
  double r = 0;
  std::string s;
};
Value vint(int v) {
  Value o;
  o.kind = 1;
  o.i = v;
  return o;
}
Value vreal(double v) {
  Value o;
  o.kind = 2;
  o.r = v;
  return o;
}
Value vstr(const std::string &v) {
  Value o;
  o.kind = 3;
  o.s = v;
  return o;
}

// Everything the writers read from one solved system.
struct State {
  scf_results r;
  std::vector<double> seg_phi, mol_phi;  // [n][M]
  std::vector<double> psi, q, eps;       // [M] each, charged systems
};

int load_state(const scf_system *S, State &st) {
  int rc = scf_get_results(S, &st.r);
  if (rc != SCF_OK) return rc;
  const size_t M = S->h.M;
  st.seg_phi.assign(S->h.nseg * M, 0.0);
  st.mol_phi.assign(S->h.nmol * M, 0.0);
  rc = scf_get_profiles(S, st.seg_phi.data(), st.mol_phi.data());
  if (rc != SCF_OK || !S->spec.charged) return rc;
  st.psi.assign(M, 0.0);
  st.q.assign(M, 0.0);
  st.eps.assign(M, 0.0);
  return scf_get_electrostatics(S, st.psi.data(), st.q.data(), st.eps.data());
}

// a system profile of a charged system (System::GetPointer src/system.cpp:1632-1650)
const double *sys_profile(const SystemSpec &s, const State &st, const OutputItem &it) {
  if (it.key != "sys" || it.name != s.sys_name || !s.charged) return nullptr;
  if (it.prop == "psi") return st.psi.data();
  if (it.prop == "q") return st.q.data();
  if (it.prop == "eps") return st.eps.data();
  return nullptr;
}

// One kal quantity (the scalars the reference's objects push for output,
// src/system.cpp PushOutput, src/molecule.cpp:1939-2075,
// src/segment.cpp:1741-1800, src/solve_scf.cpp:325-345, lattice PushOutput)
Value kal_value(const scf_system *S, const State &st, const scf_solve_info *info, const OutputItem &it) {
  const SystemSpec &s = S->spec;
  const LatticeSpec &l = s.lat;
  const std::string &p = it.prop;
  if (it.key == "sys" && it.name == s.sys_name) {
    if (p == "free_energy") return vreal(st.r.free_energy);
    if (p == "grand_potential") return vreal(st.r.grand_potential);
    if (p == "start") return vint(s.start);
    if (p == "Dpsi" && s.charged) return vreal(st.psi[l.M - 1] - st.psi[0]);
  } else if (it.key == "newton" && it.name == s.newton_name) {
    if (p == "iterations" && info) return vint(info->iterations);
    if (p == "residual") return vreal(info ? info->residual : st.r.residual);
    if (p == "tolerance") return vreal(S->h.tol);
    if (p == "iterationlimit") return vint(S->h.itlim);
    if (p == "m") return vint(s.solver.m);
    if (p == "delta_max") return vreal(s.solver.deltamax);
    if (p == "method") return vstr(s.method_name);
    if (p == "stop_criterion") return vstr(s.stop_criterion_name);
  } else if (it.key == "mol") {
    for (size_t i = 0; i < s.mol.size(); i++) {
      if (s.mol[i].name != it.name) continue;
      const MoleculeSpec &m = s.mol[i];
      if (p == "GN") return vreal(st.r.GN[i]);
      if (p == "n") return vreal(st.r.n[i]);
      if (p == "theta") return vreal(st.r.theta[i]);
      if (p == "theta_exc" || p == "thetaexc") return vreal(st.r.theta[i] - l.volume * st.r.phibulk[i]);
      if (p == "phibulk") return vreal(st.r.phibulk[i]);
      if (p == "Mu") return vreal(st.r.mu[i]);
      if (p == "chainlength") return vint(m.chainlength);
      if (p == "composition") return vstr(s.mol_composition[i]);
      if (p == "freedom") return vstr(s.mol_freedom_name[i]);
      if (p == "Markov") return vint(m.markov);
      if (p == "k_stiff") return vreal(m.k_stiff);
      if (l.gradients == 1 && (p == "phi1" || p == "phiM" || p == "Dphi")) {
        const double *ph = st.mol_phi.data() + i * (size_t)l.M;
        const double a = ph[l.fjc], b = ph[l.M - 2 * l.fjc];
        return vreal(p == "phi1" ? a : (p == "phiM" ? b : a - b));
      }
    }
  } else if (it.key == "mon") {
    for (size_t k = 0; k < s.seg.size(); k++) {
      if (s.seg[k].name != it.name) continue;
      if (p == "freedom") return vstr(s.seg_freedom_name[k]);
      if (p == "valence") return vreal(s.seg_valence[k]);
      if (p == "phibulk") return vreal(st.r.seg_phibulk[k]);
      if (p == "theta") {
        std::vector<double> X(st.seg_phi.begin() + k * (size_t)l.M, st.seg_phi.begin() + (k + 1) * (size_t)l.M);
        remove_ghosts(l, X);
        return vreal(weighted_sum(l, X));
      }
    }
  } else if (it.key == "lat") {
    if (p == "gradients") return vint(l.gradients);
    if (p == "volume") return vreal(l.volume);
    if (p == "lattice_type") return vstr(s.lattice_type_name);
    if (p == "FJC_choices") return vint(l.FJC);
    if (p == "offset_first_layer" && l.offset_first_layer > 0) return vreal(l.offset_first_layer);
    if (p == "bond_length") return vreal(l.bond_length / l.fjc);
    if (p == "n_layers") return vint(l.MX / l.fjc);
    if (p == "n_layers_y" && l.gradients >= 2) return vint(l.MY / l.fjc);
    if (p == "n_layers_z" && l.gradients == 3) return vint(l.MZ / l.fjc);
  }
  return Value();
}

std::string header_name(const OutputItem &it, const std::string &sep) { return it.key + sep + it.name + sep + it.prop; }

bool file_exists(const std::string &path) {
  std::ifstream f(path.c_str());
  return (bool)f;
}

}  // namespace

// Stem of the output files from the input file name (Output::WriteOutput,
// src/output.cpp:392-419): the part before the first '.', without the
// directories when use_output_folder is set.
std::string output_stem(const std::string &input_path, bool use_output_folder) {
  std::string stem = input_path.substr(0, input_path.find('.'));
  if (use_output_folder) {
    const size_t k = stem.find_last_of("/\\");
    if (k != std::string::npos) stem = stem.substr(k + 1);
  }
  return stem;
}

// ---- var search (see VarSearch in scf_spec.h)

// Set the searched property of the system to X (Molecule::PutValue,
// src/molecule.cpp:995-1015): theta fixes n = theta / N, n fixes theta = n N,
// phibulk the bulk volume fraction of a free molecule.
static int put_search_value(scf_system *S, double X) {
  const VarSearch &vs = S->spec.search;
  MoleculeSpec &m = S->spec.mol[vs.mol];
  if (vs.prop == "theta") {
    m.theta = X;
    m.n = X / m.chainlength;
  } else if (vs.prop == "n") {
    m.n = X;
    m.theta = X * m.chainlength;
  } else {
    m.phibulk = X;
  }
  S->h.mol_n_in[vs.mol] = m.n;
  S->h.mol_phibulk_in[vs.mol] = m.phibulk;
  return upload_solver(S);
}

// Error of the target after a solve (System::GetError src/system.cpp:
// 1405-1427, Molecule::GetError src/molecule.cpp:942-969).
static double search_error(const VarSearch &vs, const scf_results &r) {
  const double v = vs.target_value;
  if (vs.target_kind == "sys") return vs.target_prop == "free_energy" ? r.free_energy - v : -(r.grand_potential - v);
  const int i = vs.target_mol;
  if (vs.target_prop == "theta") return v != 0 ? r.theta[i] / v - 1.0 : r.theta[i];
  if (vs.target_prop == "n") return v != 0 ? r.n[i] / v - 1.0 : r.n[i];
  if (vs.target_prop == "phibulk") return -(v != 0 ? r.phibulk[i] / v - 1.0 : r.phibulk[i]);
  return v != 0 ? r.mu[i] / v - 1.0 : r.mu[i];
}

// Root of the target error over the searched value; every evaluation is a
// solve that starts from the previous one's fields.  Secant steps, each at
// most super_deltamax times the current value (and halving it rather than
// crossing zero), until the error changes sign; from then on regula falsi
// with the Illinois modification (Dowell & Jarratt, BIT 11 (1971) 168), which
// keeps the root bracketed.  Stops when |error| <= super_tolerance (*met) or
// after super_iterationlimit evaluations; `info` describes the last solve.
static int run_search(scf_system *S, scf_solve_info *info, bool *met, int *evaluations) {
  const VarSearch &vs = S->spec.search;
  *met = false;
  *evaluations = 0;
  bool solved = true;
  auto eval = [&](double X, double &e) -> int {
    int rc = put_search_value(S, X);
    if (rc == SCF_OK) rc = scf_solve_resident(&S, 1, info);
    (*evaluations)++;
    if (rc != SCF_OK) return rc;
    solved = info->status == SCF_OK;
    scf_results r;
    rc = scf_get_results(S, &r);
    e = search_error(vs, r);
    return rc;
  };
  double a = vs.start_value, fa = 0, b = 0, fb = 0;
  int rc = eval(a, fa);
  if (rc != SCF_OK || !solved) return rc;
  if (std::fabs(fa) <= vs.super_tolerance) {
    *met = true;
    return SCF_OK;
  }
  b = a != 0 ? a * 1.01 : 1e-3;
  rc = eval(b, fb);
  if (rc != SCF_OK || !solved) return rc;
  bool bracket = fa * fb < 0;
  for (int it = 0; it < vs.super_iterationlimit; it++) {
    if (std::fabs(fb) <= vs.super_tolerance) {
      *met = true;
      return SCF_OK;
    }
    if (fb == fa) return SCF_OK;  // no slope: the search stops unmet
    double c = b - fb * (b - a) / (fb - fa);
    if (!bracket) {
      const double lim = vs.super_deltamax * std::fabs(b);
      if (c > b + lim) c = b + lim;
      if (c < b - lim) c = b - lim;
      if (c <= 0) c = 0.5 * b;  // theta, n and phibulk stay positive
    }
    double fc = 0;
    rc = eval(c, fc);
    if (rc != SCF_OK || !solved) return rc;
    if (bracket) {
      if (fc * fb < 0) {
        a = b;
        fa = fb;
      } else {
        fa *= 0.5;
      }
    } else {
      a = b;
      fa = fb;
      bracket = fc * fb < 0;
    }
    b = c;
    fb = fc;
  }
  *met = std::fabs(fb) <= vs.super_tolerance;
  return SCF_OK;
}

}  // namespace nscf

using namespace nscf;

extern "C" {

int scf_write_kal(scf_system *S, const char *path, int append, const scf_solve_info *info) {
  if (!S || !path) return fail(SCF_E_ARG, "null argument");
  const OutputRequest &req = S->spec.out.kal;
  State st;
  int rc = load_state(S, st);
  if (rc != SCF_OK) return rc;
  const bool header = !(file_exists(path) && append);
  FILE *fp = std::fopen(path, header ? "w" : "a");
  if (!fp) return fail(SCF_E_ARG, std::string("cannot open ") + path);
  const size_t n = req.items.size();
  if (header) {
    for (size_t i = 0; i < n; i++)
      std::fprintf(fp, i + 1 < n ? "%s\t" : "%s", header_name(req.items[i], req.sep).c_str());
    std::fprintf(fp, "\n");
  }
  for (size_t i = 0; i < n; i++) {
    const Value v = kal_value(S, st, info, req.items[i]);
    const char *tab = i + 1 < n ? "\t" : "";
    if (v.kind == 0) std::fprintf(fp, "NiN%s", tab);
    if (v.kind == 1) std::fprintf(fp, "%i%s", v.i, tab);
    if (v.kind == 2) std::fprintf(fp, "%.16e%s", v.r, tab);
    if (v.kind == 3) std::fprintf(fp, "%s%s", v.s.c_str(), tab);
  }
  std::fprintf(fp, "\n");
  std::fclose(fp);
  return SCF_OK;
}

int scf_write_pro(scf_system *S, const char *path) {
  if (!S || !path) return fail(SCF_E_ARG, "null argument");
  const SystemSpec &s = S->spec;
  const LatticeSpec &l = s.lat;
  const OutputRequest &req = s.out.pro;
  State st;
  int rc = load_state(S, st);
  if (rc != SCF_OK) return rc;
  // the profiles the request names that exist: segment and molecule densities
  std::vector<std::string> names;
  std::vector<std::vector<double>> cols;
  for (const OutputItem &it : req.items) {
    const double *src = nullptr;
    if (it.prop == "phi" && it.key == "mon")
      for (size_t k = 0; k < s.seg.size(); k++)
        if (s.seg[k].name == it.name) src = st.seg_phi.data() + k * (size_t)l.M;
    if (it.prop == "phi" && it.key == "mol")
      for (size_t i = 0; i < s.mol.size(); i++)
        if (s.mol[i].name == it.name) src = st.mol_phi.data() + i * (size_t)l.M;
    if (!src) src = sys_profile(s, st, it);
    if (!src) continue;  // not a profile this system knows: left out, as the reference does
    std::vector<double> X(src, src + l.M);
    if (req.write_bounds) set_ghosts(l, X);
    cols.push_back(std::move(X));
    names.push_back(header_name(it, req.sep));
  }
  FILE *fp = std::fopen(path, "w");
  if (!fp) return fail(SCF_E_ARG, std::string("cannot open ") + path);
  std::fprintf(fp, l.gradients == 1 ? "x\t" : (l.gradients == 2 ? "x\ty\t" : "x\ty\tz\t"));
  for (size_t i = 0; i < names.size(); i++)
    std::fprintf(fp, i + 1 < names.size() ? "%s\t" : "%s", names[i].c_str());
  std::fprintf(fp, "\n");
  const int f = l.fjc, a = req.write_bounds ? 0 : f;
  const double off = l.offset_first_layer / f;
  auto coord = [&](int x) { return 1.0 * (x - f + 1) / f - 0.5 / f; };
  if (l.gradients == 1) {
    for (int x = a; x < l.MX + 2 * f - a; x++) {
      std::fprintf(fp, "%e\t", off + 1.0 * (x - f + 1) / f - 0.5 / f);
      for (size_t c = 0; c < cols.size(); c++)
        std::fprintf(fp, c + 1 < cols.size() ? "%.20g\t" : "%.20g", cols[c][x]);
      std::fprintf(fp, "\n");
    }
  } else if (l.gradients == 2) {
    for (int x = a; x < l.MX + 2 * f - a; x++)
      for (int y = a; y < l.MY + 2 * f - a; y++) {
        std::fprintf(fp, "%e\t%e\t", off + 1.0 * (x - f + 1) / f - 0.5 / f, coord(y));
        for (auto &col : cols) std::fprintf(fp, "%.20e\t", col[(size_t)x * l.JX + y]);
        std::fprintf(fp, "\n");
      }
  } else {
    for (int x = a; x < l.MX + 2 * f - a; x++)
      for (int y = a; y < l.MY + 2 * f - a; y++)
        for (int z = a; z < l.MZ + 2 * f - a; z++) {
          std::fprintf(fp, "%e\t%e\t%e\t", coord(x), coord(y), coord(z));
          for (auto &col : cols) std::fprintf(fp, "%.20e\t", col[(size_t)x * l.JX + (size_t)y * l.JY + z]);
          std::fprintf(fp, "\n");
        }
  }
  std::fclose(fp);
  return SCF_OK;
}

int scf_write_vec(scf_system *S, const char *path) {
  if (!S || !path) return fail(SCF_E_ARG, "null argument");
  const SystemSpec &s = S->spec;
  const LatticeSpec &l = s.lat;
  const OutputRequest &req = s.out.vec;
  State st;
  int rc = load_state(S, st);
  if (rc != SCF_OK) return rc;
  FILE *fp = std::fopen(path, "a");
  if (!fp) return fail(SCF_E_ARG, std::string("cannot open ") + path);
  // one line per requested profile, every site ghosts included: segment
  // densities with their bounds set (a frozen segment's is its mask),
  // molecule densities with their bounds set, system profiles as they are
  // (Output::WriteOutput src/output.cpp:452-473, Segment::GetPointer
  // src/segment.cpp:1870-1882, Molecule::GetPointer src/molecule.cpp:2103-2113)
  for (const OutputItem &it : req.items) {
    const double *src = nullptr;
    bool bounds = true;
    if (it.prop == "phi" && it.key == "mon")
      for (size_t k = 0; k < s.seg.size(); k++)
        if (s.seg[k].name == it.name) {
          src = st.seg_phi.data() + k * (size_t)l.M;
          bounds = s.seg[k].freedom != SEG_FROZEN;
        }
    if (it.prop == "phi" && it.key == "mol")
      for (size_t i = 0; i < s.mol.size(); i++)
        if (s.mol[i].name == it.name) src = st.mol_phi.data() + i * (size_t)l.M;
    if (!src) {
      src = sys_profile(s, st, it);
      bounds = false;
    }
    if (!src) continue;  // not a profile this system knows: left out, as the reference does
    std::vector<double> X(src, src + l.M);
    if (bounds) set_ghosts(l, X);
    std::fprintf(fp, "%s\t", header_name(it, req.sep).c_str());
    for (double v : X) std::fprintf(fp, "%e\t", v);
    std::fprintf(fp, "\n");
  }
  std::fclose(fp);
  return SCF_OK;
}

int scf_write_vtk(scf_system *S, const char *path) {
  if (!S || !path) return fail(SCF_E_ARG, "null argument");
  const SystemSpec &s = S->spec;
  const LatticeSpec &l = s.lat;
  const OutputRequest &req = s.out.vtk;
  if (l.gradients == 1 || req.items.empty()) return SCF_OK;  // the reference writes no 1-gradient vtk file
  State st;
  int rc = load_state(S, st);
  if (rc != SCF_OK) return rc;
  // one member: the first requested profile (Output::WriteOutput src/output.cpp:560-571)
  const OutputItem &it = req.items[0];
  const double *src = nullptr;
  if (it.prop == "phi" && it.key == "mon")
    for (size_t k = 0; k < s.seg.size(); k++)
      if (s.seg[k].name == it.name) src = st.seg_phi.data() + k * (size_t)l.M;
  if (it.prop == "phi" && it.key == "mol")
    for (size_t i = 0; i < s.mol.size(); i++)
      if (s.mol[i].name == it.name) src = st.mol_phi.data() + i * (size_t)l.M;
  if (!src) return SCF_OK;  // not a profile: no file, as in the reference
  std::vector<double> X(src, src + l.M);
  if (req.write_bounds) set_ghosts(l, X);
  FILE *fp = std::fopen(path, "w+");
  if (!fp) return fail(SCF_E_ARG, std::string("cannot open ") + path);
  const int f = l.fjc;
  const std::string id = header_name(it, req.sep);
  // LGrad2::vtk / LGrad3::vtk (src/LGrad2.cpp:159-183, src/LGrad3.cpp:48-79)
  if (l.gradients == 2)
    std::fprintf(fp, "# vtk DataFile Version 3.0\nvtk output\nASCII\nDATASET STRUCTURED_POINTS\nDIMENSIONS %i %i %i\n",
                 l.MY, l.MX, 1);
  else
    std::fprintf(fp, "# vtk DataFile Version 3.0\nvtk output\nASCII\nDATASET STRUCTURED_POINTS\nDIMENSIONS %i %i %i\n",
                 l.MZ, l.MY, l.MX);
  const int npts = req.write_bounds ? l.M : l.MX * l.MY * (l.gradients == 3 ? l.MZ : 1);
  std::fprintf(fp, "SPACING 1 1 1\nORIGIN 0 0 0\nPOINT_DATA %i\n", npts);
  std::fprintf(fp, "SCALARS %s double\nLOOKUP_TABLE default\n", id.c_str());
  if (l.gradients == 2) {
    // with write_bounds the 2-gradient writer prints every site and then
    // the interior again, as the reference does
    if (req.write_bounds)
      for (int i = 0; i < l.M; i++) std::fprintf(fp, "%f\n", X[i]);
    for (int x = f; x < l.MX + f; x++)
      for (int y = f; y < l.MY + f; y++) std::fprintf(fp, "%f\n", X[(size_t)x * l.JX + y]);
  } else if (req.write_bounds) {
    for (int i = 0; i < l.M; i++) std::fprintf(fp, "%f\n", X[i]);
  } else {
    for (int x = f; x < l.MX + f; x++)
      for (int y = f; y < l.MY + f; y++)
        for (int z = f; z < l.MZ + f; z++) std::fprintf(fp, "%e\n", X[(size_t)x * l.JX + (size_t)y * l.JY + z]);
  }
  std::fclose(fp);
  return SCF_OK;
}

int scf_store_guess(scf_system *S, const char *path) {
  if (!S || !path) return fail(SCF_E_ARG, "null argument");
  const SystemSpec &s = S->spec;
  const LatticeSpec &l = s.lat;
  std::vector<double> x((size_t)s.nvar());
  int rc = scf_download_fields(S, x.data());
  if (rc != SCF_OK) return rc;
  FILE *fp = std::fopen(path, "w");
  if (!fp) return fail(SCF_E_ARG, std::string("cannot open ") + path);
  std::fprintf(fp, "%s\n", s.method_name.c_str());
  std::fprintf(fp, " %i\t%i\t%i\t%i\n", l.MX, l.gradients >= 2 ? l.MY : 0, l.gradients == 3 ? l.MZ : 0, l.fjc);
  std::fprintf(fp, "%s\n", s.charged ? "true" : "false");
  std::fprintf(fp, "%i\n", (int)s.itmon.size());
  for (int k : s.itmon) std::fprintf(fp, "%s\n", s.seg[k].name.c_str());
  std::fprintf(fp, "%i\n", 0);
  for (double v : x) std::fprintf(fp, "%e\n", v);
  std::fclose(fp);
  return SCF_OK;
}

int scf_read_guess(scf_system *S, const char *path) {
  if (!S || !path) return fail(SCF_E_ARG, "null argument");
  std::ifstream f(path);
  if (!f) return fail(SCF_E_ARG, std::string("guess file ") + path + " not found");
  std::stringstream ss;
  ss << f.rdbuf();
  std::vector<double> x;
  std::string err;
  if (!read_guess_text(ss.str(), S->spec, x, err)) return fail(SCF_E_ARG, err);
  return scf_upload_fields(S, x.data());
}

int scf_run_input_file(const char *path, int *n_starts_out, int *n_converged_out) {
  if (!path) return fail(SCF_E_ARG, "null argument");
  std::ifstream f(path);
  if (!f) return fail(SCF_E_ARG, std::string("input file ") + path + " not found");
  std::stringstream ss;
  ss << f.rdbuf();
  const std::string text = ss.str();
  int n_starts = count_starts(text);
  if (n_starts < 1) n_starts = 1;
  int converged = 0, solves = 0;
  scf_system *prev = nullptr;
  int rc = SCF_OK;
  bool kal_started = false;
  for (int start = 1; start <= n_starts && rc == SCF_OK; start++) {
    // the start block as written, and its var scan (points 0 .. num_of_steps)
    SystemSpec spec0;
    std::string why;
    if (!parse_input(text, start, spec0, why)) {
      rc = fail(SCF_E_PARSE, why);
      break;
    }
    const VarScan &vs = spec0.scan;
    const int nsub = vs.present ? vs.num_of_steps : 0;
    for (int subl = 0; subl <= nsub && rc == SCF_OK; subl++) {
      std::string t = text;
      if (vs.present) {
        char num[64];
        std::snprintf(num, sizeof(num), "%.17g", scan_value(vs, subl));
        if (vs.kind == "mol") {
          t = insert_before_start(t, start, "mol : " + vs.name + " : " + vs.prop + " : " + num);
        } else if (vs.prop == "valence" || vs.prop == "ePsi0/kT") {
          t = insert_before_start(t, start, "mon : " + vs.name + " : " + (vs.prop == "valence" ? "valence" : "e.psi0/kT") +
                                                " : " + num);
        } else {
          const std::string other = vs.prop.substr(4);
          t = insert_before_start(t, start, "mon : " + vs.name + " : chi_" + other + " : " + num);
          t = insert_before_start(t, start, "mon : " + other + " : chi_" + vs.name + " : " + num);
        }
      }
      scf_system *S = nullptr;
      char err[512] = {0};
      rc = scf_system_create(t.c_str(), start, &S, err, sizeof(err));
      if (rc != SCF_OK) break;
      const SystemSpec &s = S->spec;
      // initial guess (namics main loop, src/namics.cpp:419-466,530-570):
      // the guess file is read once per start block; every later solve
      // starts from the one before it
      if (s.initial_guess == "file" && subl == 0) {
        rc = scf_read_guess(S, s.guess_inputfile.c_str());
      } else if (s.initial_guess != "none" && prev) {
        rc = scf_carry_fields(S, prev);
      }
      scf_solve_info info{};
      if (rc == SCF_OK && s.search.present) {
        // var search: solves until the target is met; counted as one solve
        bool met = false;
        int evals = 0;
        rc = run_search(S, &info, &met, &evals);
        if (rc == SCF_OK && info.status == SCF_OK && !met) info.status = SCF_NOT_CONVERGED;
      } else if (rc == SCF_OK) {
        rc = scf_solve_resident(&S, 1, &info);
      }
      solves++;
      if (rc == SCF_OK && info.status == SCF_OK) converged++;
      const std::string stem = output_stem(path, s.out.use_output_folder);
      const std::string tag = (n_starts > 1 ? "_" + std::to_string(start) : std::string()) +
                              (subl > 0 ? "_" + std::to_string(subl) : std::string());
      if (rc == SCF_OK && s.out.kal.present && s.out.kal.write) {
        // the first kal row of a run appends only when asked (or when the
        // request sets no append option); every later row appends
        const bool append = kal_started || s.out.kal.append || !s.out.kal.append_given;
        rc = scf_write_kal(S, (s.out.path + stem + ".kal").c_str(), append ? 1 : 0, &info);
        kal_started = true;
      }
      if (rc == SCF_OK && s.out.pro.present && s.out.pro.write)
        rc = scf_write_pro(S, (s.out.path + stem + tag + ".pro").c_str());
      if (rc == SCF_OK && s.out.vtk.present && s.out.vtk.write)
        rc = scf_write_vtk(S, (s.out.path + stem + tag + ".vtk").c_str());
      if (rc == SCF_OK && s.out.vec.present && s.out.vec.write)
        rc = scf_write_vec(S, (s.out.path + stem + ".vec").c_str());
      if (rc == SCF_OK && s.out.pos.present && s.out.pos.write) {
        // pos prints integer arrays, and no object offers one
        // (Segment/Molecule/System::GetPointerInt return none,
        // src/segment.cpp:1922-1930): the file is opened for appending and
        // stays as it was (src/output.cpp:429-450)
        FILE *fp = std::fopen((s.out.path + stem + ".pos").c_str(), "a");
        if (!fp) rc = fail(SCF_E_ARG, "cannot open " + s.out.path + stem + ".pos");
        else std::fclose(fp);
      }
      if (rc == SCF_OK && s.final_guess == "file") {
        std::string name;
        const std::string gtag = "_" + std::to_string(start) + (subl > 0 ? "_" + std::to_string(subl) : std::string());
        if (s.guess_outputfile.empty()) {
          name = stem + gtag + ".outiv";
        } else {
          const std::string &g = s.guess_outputfile;
          const size_t d = g.find('.');
          const std::string ext = d == std::string::npos ? "" : g.substr(d + 1, g.find('.', d + 1) - d - 1);
          name = g.substr(0, d) + gtag + "." + ext;
        }
        rc = scf_store_guess(S, (s.out.path + name).c_str());
      }
      if (prev) scf_system_destroy(prev);
      prev = S;
    }
  }
  if (prev) scf_system_destroy(prev);
  if (n_starts_out) *n_starts_out = solves;
  if (n_converged_out) *n_converged_out = converged;
  return rc;
}

}  // extern "C"
