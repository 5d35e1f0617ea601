// `.in` reader and system builder for the SF-SCF hot path.  See scf_spec.h.
//
// Grammar (reference src/input.cpp:3-78): every line is `key : name : prop :
// value`; all spaces and tabs are removed; lines starting with `//` are
// comments; a line `start` closes a block and the values of block k are the
// union of blocks 1..k with the last value winning (src/input.cpp:366-429).
// Lists of names (mon, mol) keep the order of first appearance
// (src/input.cpp:286-306).
#include "scf_spec.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <map>
#include <sstream>

namespace nscf {
namespace {

struct Line {
  std::string key, name, prop, value;
};

struct Parsed {
  std::vector<Line> lines;                 // settings visible at the chosen start
  std::vector<std::string> mons, mols;     // order of first appearance
  std::string lat_name, sys_name = "NN", newton_name;
};

std::vector<std::string> split(const std::string &s, char c) {
  std::vector<std::string> out;
  std::string cur;
  for (char ch : s) {
    if (ch == c) {
      out.push_back(cur);
      cur.clear();
    } else {
      cur.push_back(ch);
    }
  }
  out.push_back(cur);
  return out;
}

std::vector<std::string> clean_lines(const std::string &text) {
  std::vector<std::string> out;
  std::istringstream is(text);
  std::string line;
  while (std::getline(is, line)) {
    std::string t;
    for (char ch : line)
      if (ch != ' ' && ch != '\t' && ch != '\r') t.push_back(ch);
    if (t.empty()) continue;
    if (t.size() >= 2 && t.compare(0, 2, "//") == 0) continue;
    out.push_back(t);
  }
  if (out.empty() || out.back().compare(0, 5, "start") != 0) out.push_back("start");
  return out;
}

bool collect(const std::string &text, int start, Parsed &p, std::string &err) {
  auto lines = clean_lines(text);
  int n_start = 0;
  for (auto &l : lines) {
    if (l.compare(0, 5, "start") == 0) {
      n_start++;
      if (n_start >= start) break;
      continue;
    }
    auto f = split(l, ':');
    if (f.size() < 4) {
      err = "input line '" + l + "' does not have the form key : name : prop : value";
      return false;
    }
    Line ln;
    ln.key = f[0];
    ln.name = f[1];
    ln.prop = f[2];
    ln.value = f.size() >= 4 ? f[3] : "";
    if (f.size() > 4) {  // values may contain ':' only in exotic keys; keep the rest
      for (size_t i = 4; i < f.size(); i++) ln.value += ":" + f[i];
    }
    if (ln.key == "mon" && std::find(p.mons.begin(), p.mons.end(), ln.name) == p.mons.end())
      p.mons.push_back(ln.name);
    if (ln.key == "mol" && std::find(p.mols.begin(), p.mols.end(), ln.name) == p.mols.end())
      p.mols.push_back(ln.name);
    if (ln.key == "lat") p.lat_name = ln.name;
    if (ln.key == "sys") p.sys_name = ln.name;
    if (ln.key == "newton") p.newton_name = ln.name;
    p.lines.push_back(ln);
  }
  if (n_start < start && start > 1) {
    err = "input has fewer than the requested number of start blocks";
    return false;
  }
  return true;
}

// last value wins
bool get(const Parsed &p, const std::string &key, const std::string &name, const std::string &prop,
         std::string &val) {
  bool found = false;
  for (auto &l : p.lines)
    if (l.key == key && l.name == name && l.prop == prop) {
      val = l.value;
      found = true;
    }
  return found;
}

bool to_int(const std::string &s, int &v) {
  if (s.empty()) return false;
  char *end = nullptr;
  long r = std::strtol(s.c_str(), &end, 10);
  if (*end != 0) return false;
  v = (int)r;
  return true;
}

bool to_real(const std::string &s, double &v) {
  if (s.empty()) return false;
  char *end = nullptr;
  double r = std::strtod(s.c_str(), &end);
  if (*end != 0) return false;
  v = r;
  return true;
}

bool to_bool(const std::string &s, bool &v) {
  if (s == "true" || s == "TRUE" || s == "True" || s == "1") {
    v = true;
    return true;
  }
  if (s == "false" || s == "FALSE" || s == "False" || s == "0") {
    v = false;
    return true;
  }
  return false;
}

int bc_of(const std::string &s, bool &ok) {
  ok = true;
  if (s == "mirror") return BC_MIRROR;
  if (s == "periodic") return BC_PERIODIC;
  if (s == "surface") return BC_SURFACE;
  ok = false;
  return BC_MIRROR;
}

// Boundary index tables for fjc == 1 (src/lattice.cpp:94-196).
void boundary_indices(LatticeSpec &l) {
  auto lower = [](int bc, int Mdim) { return bc == BC_MIRROR ? 1 : (bc == BC_PERIODIC ? Mdim : 0); };
  auto upper = [](int bc, int Mdim) { return bc == BC_MIRROR ? Mdim : (bc == BC_PERIODIC ? 1 : Mdim + 1); };
  l.BX1 = lower(l.bc[0], l.MX);
  l.BXM = upper(l.bc[3], l.MX);
  l.BY1 = lower(l.bc[1], l.MY);
  l.BYM = upper(l.bc[4], l.MY);
  l.BZ1 = lower(l.bc[2], l.MZ);
  l.BZM = upper(l.bc[5], l.MZ);
  // refined 1-gradient lattice: fjc ghost layers per side
  const int f = l.fjc;
  l.bx1f.assign(f, 0);
  l.bxmf.assign(f, 0);
  for (int k = 0; k < f; k++) {
    l.bx1f[k] = l.bc[0] == BC_MIRROR ? 2 * f - 1 - k : (l.bc[0] == BC_PERIODIC ? l.MX + k : k);
    l.bxmf[k] = l.bc[3] == BC_MIRROR ? l.MX + f - k - 1 : (l.bc[3] == BC_PERIODIC ? f + k : l.MX + f + k);
  }
  l.by1f.assign(f, 0);
  l.bymf.assign(f, 0);
  for (int k = 0; k < f; k++) {
    l.by1f[k] = l.bc[1] == BC_MIRROR ? 2 * f - 1 - k : (l.bc[1] == BC_PERIODIC ? l.MY + k : k);
    l.bymf[k] = l.bc[4] == BC_MIRROR ? l.MY + f - k - 1 : (l.bc[4] == BC_PERIODIC ? f + k : l.MY + f + k);
  }
  l.bz1f.assign(f, 0);
  l.bzmf.assign(f, 0);
  for (int k = 0; k < f; k++) {
    l.bz1f[k] = l.bc[2] == BC_MIRROR ? 2 * f - 1 - k : (l.bc[2] == BC_PERIODIC ? l.MZ + k : k);
    l.bzmf[k] = l.bc[5] == BC_MIRROR ? l.MZ + f - k - 1 : (l.bc[5] == BC_PERIODIC ? f + k : l.MZ + f + k);
  }
}

// Weights of the refined (fjc > 1) 2-gradient lattices (LGrad2::
// ComputeLambdas src/LGrad2.cpp:51-97, LG2Planar::ComputeLambdas
// src/LG2Planar.cpp:15-24).  Cylindrical: the radial rows of the 1-gradient
// construction at every interior site (x, y), site volume L = 2 PIE r / fjc;
// planar: no LAMBDA rows (the planar stencil is a plain block sum) and
// L = 1/fjc for fjc = 2 only, as in the reference.
void refined_lambdas2(LatticeSpec &l) {
  const int M = l.M, f = l.fjc, F = l.FJC, MX = l.MX, MY = l.MY, JX = l.JX;
  const double off = l.offset_first_layer;
  l.L.assign(M, 0.0);
  l.lambda_1.assign(M, 0.0);
  l.lambda0.assign(M, 0.0);
  l.lambda1.assign(M, 0.0);
  l.LAMBDA.assign((size_t)F * M, 0.0);
  if (l.geometry == GEOM_PLANAR) {
    if (f == 2)
      for (int i = 0; i < M; i++) l.L[i] = 1.0 / f;
    return;
  }
  double *LA = l.LAMBDA.data();
  const double c = 1.0 / (1.0 * F - 1.0), top = 1.0 * MX / f;
  for (int y = f; y < MY + f; y++)
    for (int x = f; x < MX + f; x++) {
      const int i = x * JX + y;
      const double r = off + 1.0 * (x - f + 1.0) / f;
      double rlow = r - 0.5, rhigh = r + 0.5;
      l.L[i] = kPIE * (2.0 * r) / f;
      const double VL = l.L[i] / kPIE * f;
      if ((rlow - r) * 2 + r > 0.0) LA[i] += c * rlow / VL;
      if ((rhigh - r) * 2 + r < top) {
        LA[i + (F - 1) * M] += c * rhigh / VL;
      } else {
        const double d = 2 * rhigh - r - top;
        if (d > -0.001 && d < 0.001) LA[i + (F - 1) * M] += c * rhigh / VL;
        for (int j = 1; j <= f; j++)
          if (d > 0.99 * j / f && d < 1.01 * j / f) LA[i + (F - 1) * M] += c * (rhigh - 1.0 * j / f) / VL;
      }
      for (int j = 1; j < f; j++) {
        rlow += 0.5 / (f);
        rhigh -= 0.5 / (f);
        if ((rlow - r) * 2 + r > 0.0) LA[i + j * M] += c * 2.0 * rlow / VL;
        if ((rhigh - r) * 2 + r < off + top) {
          LA[i + (F - 1 - j) * M] += c * 2.0 * rhigh / VL;
        } else {
          const double d = 2 * rhigh - r - top;
          if (d > -0.001 && d < 0.001) LA[i + (F - 1 - j) * M] += c * 2.0 * rhigh / VL;
          for (int k = 1; k <= f; k++)
            if (d > 0.99 * k / f && d < 1.01 * k / f) LA[i + (F - 1 - j) * M] += c * 2.0 * (rhigh - 1.0 * k / f) / VL;
        }
      }
      double LS = 0;
      for (int j = 0; j < F; j++) LS += LA[i + j * M];
      LA[i + (F / 2) * M] += 1.0 - LS;
    }
}

// Weights of the refined (fjc > 1) 1-gradient lattices: LG1Planar::
// ComputeLambdas (src/LG1Planar.cpp:14-25) and the fjc > 1 branch of
// LGrad1::ComputeLambdas (src/LGrad1.cpp:72-163).  The curved weights split
// the bond directions over the FJC sub-layers by the radii of their shells;
// the expressions keep the reference's operation order so that the weights
// are the same doubles.
void refined_lambdas(LatticeSpec &l) {
  const int M = l.M, f = l.fjc, F = l.FJC, MX = l.MX;
  const double off = l.offset_first_layer;
  if (l.gradients == 2) {
    refined_lambdas2(l);
    return;
  }
  l.L.assign(M, 0.0);
  l.lambda_1.assign(M, 0.0);  // the fjc = 1 weights stay zero (unused)
  l.lambda0.assign(M, 0.0);
  l.lambda1.assign(M, 0.0);
  l.LAMBDA.assign((size_t)F * M, 0.0);
  double *LA = l.LAMBDA.data();
  if (l.geometry == GEOM_PLANAR) {
    for (int i = 0; i < M; i++) {
      l.L[i] = 1.0 / f;
      LA[i] = 1.0 / (2 * (F - 1));
      LA[i + (F - 1) * M] = 1.0 / (2 * (F - 1));
      LA[i + (F - 1) / 2 * M] = 1.0 / (F - 1);
      for (int j = 1; j < F / 2; j++) {
        LA[i + j * M] = 1.0 / (F - 1);
        LA[i + (F - j - 1) * M] = 1.0 / (F - 1);
      }
    }
    return;
  }
  const bool cyl = l.geometry == GEOM_CYLINDRICAL;
  const double c = 1.0 / (1.0 * F - 1.0), top = 1.0 * MX / f;
  // shell weight of a sub-layer boundary at radius q (outer / inner rows)
  auto wo = [&](double q, double VL) { return cyl ? c * q / VL : 0.5 / (1.0 * F - 1.0) * 4.0 * q * q / VL; };
  auto wi = [&](double q, double VL) { return cyl ? c * 2.0 * q / VL : c * 4.0 * q * q / VL; };
  for (int i = f; i < M - f; i++) {
    const double r = cyl ? off + 1.0 * (i - f + 1.0) / f : off + 1.0 * (1.0 * i - 1.0 * f + 1.0) / f;
    double rlow = r - 0.5, rhigh = r + 0.5;
    l.L[i] = cyl ? kPIE * (2.0 * r) / f : kPIE * 4.0 / 3.0 * (rhigh * rhigh * rhigh - rlow * rlow * rlow) / f;
    const double VL = l.L[i] / kPIE * f;
    if ((rlow - r) * 2 + r > 0.0) LA[i] += wo(rlow, VL);
    if ((rhigh - r) * 2 + r < top) {
      LA[i + (F - 1) * M] += wo(rhigh, VL);
    } else {
      const double d = 2 * rhigh - r - top;
      if (d > -0.001 && d < 0.001) LA[i + (F - 1) * M] += wo(rhigh, VL);
      for (int j = 1; j <= f; j++)
        if (d > 0.99 * j / f && d < 1.01 * j / f) LA[i + (F - 1) * M] += wo(rhigh - 1.0 * j / f, VL);
    }
    for (int j = 1; j < f; j++) {
      rlow += 0.5 / (f);
      rhigh -= 0.5 / (f);
      if ((rlow - r) * 2 + r > 0.0) LA[i + j * M] += wi(rlow, VL);
      if ((rhigh - r) * 2 + r < off + top) {
        LA[i + (F - 1 - j) * M] += wi(rhigh, VL);
      } else {
        const double d = 2 * rhigh - r - top;
        if (d > -0.001 && d < 0.001) LA[i + (F - 1 - j) * M] += wi(rhigh, VL);
        for (int k = 1; k <= f; k++)
          if (d > 0.99 * k / f && d < 1.01 * k / f) LA[i + (F - 1 - j) * M] += wi(rhigh - 1.0 * k / f, VL);
      }
    }
    double LS = 0;
    for (int j = 0; j < F; j++) LS += LA[i + j * M];
    LA[i + (F / 2) * M] += 1.0 - LS;
  }
}

// Site volumes and lambda weights (LG1Planar::ComputeLambdas
// src/LG1Planar.cpp:11-26, LGrad1::ComputeLambdas src/LGrad1.cpp:16-68,
// LGrad2::ComputeLambdas src/LGrad2.cpp:13-49).
void compute_lambdas(LatticeSpec &l, bool markov2) {
  const int M = l.M;
  if (l.gradients == 3) return;
  if (l.fjc > 1) {
    refined_lambdas(l);
    return;
  }
  l.L.assign(M, 0.0);
  l.lambda_1.assign(M, 0.0);
  l.lambda0.assign(M, 0.0);
  l.lambda1.assign(M, 0.0);
  const double lam = l.lambda;
  if (l.gradients == 1) {
    if (l.geometry == GEOM_PLANAR) {
      for (int i = 1; i < l.MX + 1; i++) l.L[i] = 1.0;
    } else if (l.geometry == GEOM_CYLINDRICAL) {
      for (int i = 1; i < l.MX + 1; i++) {
        double r = l.offset_first_layer + i;
        l.L[i] = kPIE * (std::pow(r, 2) - std::pow(r - 1, 2));
        l.lambda1[i] = 2.0 * kPIE * r / l.L[i] * lam;
        l.lambda_1[i] = 2.0 * kPIE * (r - 1) / l.L[i] * lam;
        l.lambda0[i] = 1.0 - 2.0 * lam;
      }
    } else {
      for (int i = 1; i < l.MX + 1; i++) {
        double r = l.offset_first_layer + i;
        l.L[i] = 4.0 / 3.0 * kPIE * (std::pow(r, 3) - std::pow(r - 1, 3));
        l.lambda1[i] = 4.0 * kPIE * std::pow(r, 2) / l.L[i] * lam;
        l.lambda_1[i] = 4.0 * kPIE * std::pow(r - 1, 2) / l.L[i] * lam;
        l.lambda0[i] = 1.0 - l.lambda1[i] - l.lambda_1[i];
      }
    }
    if (markov2 && l.geometry != GEOM_PLANAR) {
      l.l1.assign(M, 0.0);
      l.l11.assign(M, 0.0);
      l.l_1.assign(M, 0.0);
      l.l_11.assign(M, 0.0);
      for (int i = 1; i < l.MX + 1; i++) {
        l.l1[i] = l.lambda1[i] / lam;
        l.l11[i] = 1.0 - l.l1[i];
        l.l_1[i] = l.lambda_1[i] / lam;
        l.l_11[i] = 1.0 - l.l_1[i];
      }
    }
  } else {  // 2 gradients
    if (l.geometry == GEOM_CYLINDRICAL) {
      for (int x = 1; x < l.MX + 1; x++)
        for (int y = 1; y < l.MY + 1; y++) {
          int p = x * l.JX + y;
          double r = l.offset_first_layer + 1.0 * x;
          l.L[p] = kPIE * (std::pow(r, 2) - std::pow(r - 1, 2));
          l.lambda1[p] = 2.0 * kPIE * r / l.L[p] * lam;
          l.lambda_1[p] = 2.0 * kPIE * (r - 1) / l.L[p] * lam;
          l.lambda0[p] = 1.0 - 2.0 * lam;
        }
      if (markov2) {
        // Markov-2 reflection weights at every site (src/LGrad2.cpp:42-47)
        l.l1.assign(M, 0.0);
        l.l11.assign(M, 0.0);
        l.l_1.assign(M, 0.0);
        l.l_11.assign(M, 0.0);
        for (int i = 0; i < M; i++) {
          l.l1[i] = l.lambda1[i] / lam;
          l.l11[i] = 1.0 - l.l1[i];
          l.l_1[i] = l.lambda_1[i] / lam;
          l.l_11[i] = 1.0 - l.l_1[i];
        }
      }
    } else {
      // planar 2-D: L=1 on every site (LG2Planar::ComputeLambdas src/LG2Planar.cpp:15-20)
      for (int i = 0; i < M; i++) l.L[i] = 1.0;
    }
  }
}

// Parse a coordinate token of a range (firstlayer / lastlayer / lowerbound /
// upperbound / number) for dimension d.
bool coord(const LatticeSpec &l, const std::string &tok, int d, int &v) {
  // coordinates count bond-length layers: n_layers = MX / fjc
  int n = d == 0 ? l.MX / l.fjc : (d == 1 ? l.MY / l.fjc : l.MZ / l.fjc);
  if (tok == "firstlayer") {
    v = 1;
    return true;
  }
  if (tok == "lastlayer") {
    v = n;
    return true;
  }
  if (tok == "lowerbound") {
    v = 0;
    return l.bc[d] == BC_SURFACE;
  }
  if (tok == "upperbound") {
    v = n + 1;
    return l.bc[d + 3] == BC_SURFACE;
  }
  return to_int(tok, v);
}

// frozen_range / pinned_range (Segment::ParseFreedoms src/segment.cpp:230-560,
// LGrad1/LGrad2/LGrad3::ReadRange + CreateMASK): a block of sites given by
// two corner points, or one of the keywords lowerbound[_x|_y|_z],
// upperbound[_x|_y|_z], firstlayer[_x...], lastlayer[_x...].
bool parse_range(const LatticeSpec &l, std::string r, std::vector<int> &mask, std::string &err) {
  const int g = l.gradients;
  auto whole = [&](const std::string &which, int d) {
    // build "a;b" where dimension d is fixed to which and others span 1..n
    std::string lo, hi;
    for (int k = 0; k < g; k++) {
      int n = (k == 0 ? l.MX : (k == 1 ? l.MY : l.MZ)) / l.fjc;
      std::string a = k == d ? which : "1";
      std::string b = k == d ? which : std::to_string(n);
      lo += (k ? "," : "") + a;
      hi += (k ? "," : "") + b;
    }
    return lo + ";" + hi;
  };
  static const char *axes[3] = {"_x", "_y", "_z"};
  for (const char *kw : {"lowerbound", "upperbound", "firstlayer", "lastlayer"}) {
    if (r == kw) r = whole(kw, 0);
    for (int d = 0; d < 3; d++)
      if (r == std::string(kw) + axes[d]) r = whole(kw, d);
  }
  auto pts = split(r, ';');
  if (pts.size() != 2) {
    err = "range '" + r + "' must have the form x1[,y1[,z1]];x2[,y2[,z2]]";
    return false;
  }
  int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  for (int k = 0; k < 2; k++) {
    auto c = split(pts[k], ',');
    if ((int)c.size() != g) {
      err = "range '" + r + "' needs " + std::to_string(g) + " coordinates per point";
      return false;
    }
    for (int d = 0; d < g; d++) {
      int v;
      if (!coord(l, c[d], d, v)) {
        err = "range coordinate '" + c[d] + "' not understood (a bound needs a surface boundary)";
        return false;
      }
      (k == 0 ? lo : hi)[d] = v;
    }
  }
  for (int d = 0; d < g; d++)
    if (lo[d] > hi[d]) {
      err = "range '" + r + "' has its first point above its second";
      return false;
    }
  // a bond-length layer spans fjc grid layers (src/segment.cpp:145-146)
  for (int d = 0; d < g; d++) {
    lo[d] *= l.fjc;
    hi[d] = (hi[d] + 1) * l.fjc - 1;
  }
  for (int x = lo[0]; x <= hi[0]; x++)
    for (int y = (g > 1 ? lo[1] : 0); y <= (g > 1 ? hi[1] : 0); y++)
      for (int z = (g > 2 ? lo[2] : 0); z <= (g > 2 ? hi[2] : 0); z++) mask[site(l, x, y, z)] = 1;
  return true;
}

// "(A)10(B)10" -> blocks (linear chains only)
bool parse_composition(const std::string &c, const std::vector<std::string> &mons,
                       std::vector<BlockSpec> &blocks, std::string &err) {
  size_t i = 0;
  while (i < c.size()) {
    if (c[i] != '(') {
      err = "composition '" + c + "': only block sequences like (A)10(B)10, with side groups in square brackets, "
            "are supported (no aliases, repeat groups or @ keywords)";
      return false;
    }
    size_t close = c.find(')', i);
    if (close == std::string::npos) {
      err = "composition '" + c + "': unbalanced bracket";
      return false;
    }
    std::string name = c.substr(i + 1, close - i - 1);
    if (name.find('(') != std::string::npos || name.find('[') != std::string::npos) {
      err = "composition '" + c + "': branched architectures are not supported";
      return false;
    }
    size_t j = close + 1;
    while (j < c.size() && isdigit((unsigned char)c[j])) j++;
    int n = 0;
    if (!to_int(c.substr(close + 1, j - close - 1), n) || n < 1) {
      err = "composition '" + c + "': block length must be a positive integer";
      return false;
    }
    auto it = std::find(mons.begin(), mons.end(), name);
    if (it == mons.end()) {
      err = "composition '" + c + "': segment '" + name + "' is not defined";
      return false;
    }
    blocks.push_back({(int)(it - mons.begin()), n});
    i = j;
  }
  if (blocks.empty()) {
    err = "empty composition";
    return false;
  }
  return true;
}

// ---- branched compositions
// Positions of '[' and of ']' in order of appearance (two separate lists,
// not matched pairs), as Input::EvenSquareBrackets (src/input.cpp:89-100)
// returns them; false when the brackets do not balance.
bool square_positions(const std::string &c, std::vector<int> &open, std::vector<int> &close) {
  int depth = 0;
  for (int i = 0; i < (int)c.size(); i++) {
    if (c[i] == '[') {
      open.push_back(i);
      depth++;
    } else if (c[i] == ']') {
      close.push_back(i);
      if (depth-- == 0) return false;
    }
  }
  return depth == 0;
}

// The integer a repeat count starts with (Input::Get_int src/input.cpp:162-174:
// a leading digit, or '-' and a digit; anything else gives `dflt`).
int leading_int(const std::string &t, int dflt) {
  size_t i = 0;
  bool neg = false;
  if (t.size() > 1 && t[0] == '-' && isdigit((unsigned char)t[1])) {
    neg = true;
    i = 1;
  } else if (t.empty() || !isdigit((unsigned char)t[0])) {
    return dflt;
  }
  long v = 0;
  while (i < t.size() && isdigit((unsigned char)t[i]) && v < 100000000) v = v * 10 + (t[i++] - '0');
  return (int)(neg ? -v : v);
}

// Splits a branched composition into blocks with their generations, as the
// reference's Molecule::GenerateTree / Interpret (src/molecule.cpp:1155-1280)
// do: generation numbers come from the order of the brackets, and the block
// in front of a deeper generation gives its last segment to a block of its
// own (the branch point is one segment long).
struct BranchParser {
  const std::string &c;
  const std::vector<std::string> &mons;
  MoleculeSpec &m;
  std::string &err;
  std::vector<int> open, close;
  int chainlength = 0;

  void add_generation() {
    m.first_s.push_back(-1);
    m.last_s.push_back(-1);
    m.first_b.push_back(-1);
    m.last_b.push_back(-1);
  }

  bool piece(const std::string &t, int gen) {
    if (t == "[") return true;
    if (t.find(':') != std::string::npos) {
      err = "composition '" + c + "': aliases are not supported";
      return false;
    }
    std::vector<int> po, pc;
    int depth = 0;
    for (int i = 0; i < (int)t.size(); i++) {
      if (t[i] == '(') {
        po.push_back(i);
        depth++;
      } else if (t[i] == ')') {
        pc.push_back(i);
        if (depth-- == 0) break;
      }
    }
    if (depth != 0 || po.empty() || po.size() != pc.size()) {
      err = "composition '" + c + "': piece '" + t + "' is not a sequence of (segment)n";
      return false;
    }
    for (size_t k = 0; k < po.size(); k++) {
      const std::string name = t.substr(po[k] + 1, pc[k] - po[k] - 1);
      auto it = std::find(mons.begin(), mons.end(), name);
      if (it == mons.end()) {
        err = "composition '" + c + "': segment '" + name + "' is not defined";
        return false;
      }
      const size_t L = m.blocks.size();
      if (L > 0 && m.gnr[L - 1] < gen && m.blocks[L - 1].n > 1) {
        m.blocks[L - 1].n--;
        m.blocks.push_back({m.blocks[L - 1].seg, 1});
        m.gnr.push_back(m.gnr[L - 1]);
        m.last_b[m.gnr[L - 1]]++;
      }
      m.blocks.push_back({(int)(it - mons.begin()), 0});
      m.gnr.push_back(gen);
      const int nb = (int)m.blocks.size() - 1;
      if (m.first_s[gen] < 0) m.first_s[gen] = chainlength;
      if (m.first_b[gen] < 0) m.first_b[gen] = nb;
      m.last_b[gen] = nb;
      const int n = leading_int(t.substr(pc[k] + 1), 0);
      if (n < 1) {
        err = "composition '" + c + "': the number of repeats must be a positive integer";
        return false;
      }
      m.blocks.back().n = n;
      chainlength += n;
      m.last_s[gen] = chainlength;
    }
    return true;
  }

  bool tree(int gen, int &pos) {
    const int len = (int)c.size(), L = (int)open.size();
    int pos_open = 0, pos_close = len;
    while (pos_open < pos_close) {
      pos_open = len + 1;
      pos_close = len;
      bool of = false, cf = false;
      int gen_open = 0, gen_close = 0;
      for (int i = 0; i < L && !(of && cf); i++) {
        if (close[i] > pos && !cf) {
          cf = true;
          pos_close = close[i] + 1;
          gen_close = i + 1;
        }
        if (open[i] >= pos && !of) {
          of = true;
          pos_open = open[i] + 1;
          gen_open = i + 1;
        }
      }
      if (pos_close < pos_open) {
        const std::string t = c.substr(pos, pos_close - pos);
        if (!t.empty() && t[0] == '[') {
          pos = pos + 1;
          add_generation();
          if (!tree(gen_close, pos)) return false;
          pos_close = pos_open + 1;
        } else {
          pos = pos_close;
          if (!piece(t, gen)) return false;
        }
      } else {
        const std::string t = c.substr(pos, pos_open - pos);
        pos = pos_open;
        if (!piece(t, gen)) return false;
        add_generation();
        if (!tree(gen_open, pos)) return false;
      }
    }
    return true;
  }
};

// A symmetric dendrimer, `@dend(A,(B)2,3; A,(B)2,3)`: per generation a
// branch-point segment, a spacer and a functionality (Molecule::Decomposition
// src/molecule.cpp:1437-1555).  The blocks run along one path from the core
// (branch point, spacer, next branch point, ...); a block's degeneracy is
// the number of copies of it in the molecule.
bool parse_dendrimer(const std::string &c, const std::vector<std::string> &mons, MoleculeSpec &m,
                     std::string &err) {
  if (c.size() < 7 || c.compare(0, 6, "@dend(") != 0 || c.back() != ')') {
    err = "composition '" + c + "': a dendrimer is written @dend(A,(B)n,f; ...)";
    return false;
  }
  const std::string s = c.substr(6, c.size() - 7);
  if (s.find(':') != std::string::npos || s.find('#') != std::string::npos) {
    err = "composition '" + c + "': aliases are not supported";
    return false;
  }
  m.blocks.clear();
  m.block_deg.clear();
  m.n_arm.clear();
  m.first_a.clear();
  m.last_a.clear();
  m.first_s.clear();
  m.last_s.clear();
  m.first_b.clear();
  m.last_b.clear();
  m.asymmetric = false;
  int degeneracy = 1, N = -1, arm = -1;
  for (const std::string &gen : split(s, ';')) {
    auto sub = split(gen, ',');
    if (sub.size() < 3 || (sub.size() - 1) % 2 != 0) {
      err = "composition '" + c + "': each generation is branch point, then spacer and functionality per arm type";
      return false;
    }
    if (sub.size() > 3) m.asymmetric = true;
    auto it = std::find(mons.begin(), mons.end(), sub[0]);
    if (it == mons.end()) {
      err = "composition '" + c + "': branch point '" + sub[0] + "' is not a defined segment";
      return false;
    }
    m.blocks.push_back({(int)(it - mons.begin()), 1});
    m.block_deg.push_back(degeneracy);
    N++;
    m.first_a.push_back(arm + 1);
    int arms = 0;
    for (size_t k = 1; k + 1 < sub.size(); k += 2) {
      arm++;
      const int f = leading_int(sub[k + 1], 0);
      if (f < 1) {
        err = "composition '" + c + "': the number of arms of an arm type must be a positive integer";
        return false;
      }
      std::vector<BlockSpec> spacer;
      if (!parse_composition(sub[k], mons, spacer, err)) return false;
      m.n_arm.push_back(f);
      arms += f;
      m.first_s.push_back(N + 1);
      m.first_b.push_back((int)m.blocks.size());
      for (const BlockSpec &b : spacer) {
        m.blocks.push_back(b);
        N += b.n;
        m.block_deg.push_back(degeneracy * f);
      }
      m.last_b.push_back((int)m.blocks.size() - 1);
      m.last_s.push_back(N);
    }
    m.last_a.push_back(arm);
    degeneracy *= arms;
  }
  m.dendrimer = true;
  return true;
}

bool parse_comb(const std::string &c, const std::vector<std::string> &mons, MoleculeSpec &m, std::string &err) {
  if (c.size() < 7 || c.compare(0, 6, "@comb(") != 0 || c.back() != ')') {
    err = "composition '" + c + "': a comb is written @comb((A)5; A,(B)3,(A)6,4; (A)7)";
    return false;
  }
  const std::string s = c.substr(6, c.size() - 7);
  if (s.find(':') != std::string::npos || s.find('#') != std::string::npos) {
    err = "composition '" + c + "': aliases are not supported";
    return false;
  }
  auto gens = split(s, ';');
  if (gens.size() != 3) {
    err = "composition '" + c + "': a comb has a head, a repeat and a tail separated by ';'";
    return false;
  }
  auto mid = split(gens[1], ',');
  if (mid.size() != 4) {
    err = "composition '" + c + "': the repeat of a comb is branch point, tooth, spacer, number of teeth";
    return false;
  }
  const int narm = leading_int(mid[3], 0);
  if (narm < 1) {
    err = "composition '" + c + "': the number of teeth must be a positive integer";
    return false;
  }
  auto bp = std::find(mons.begin(), mons.end(), mid[0]);
  if (bp == mons.end()) {
    err = "composition '" + c + "': branch point '" + mid[0] + "' is not a defined segment";
    return false;
  }
  m.blocks.clear();
  m.block_deg.clear();
  m.n_arm.assign(1, narm);
  m.first_s.clear();
  m.last_s.clear();
  m.first_b.clear();
  m.last_b.clear();
  int N = -1;
  auto generation = [&](const std::string &comp, int deg) {
    std::vector<BlockSpec> bl;
    if (!parse_composition(comp, mons, bl, err)) return false;
    m.first_s.push_back(N + 1);
    m.first_b.push_back((int)m.blocks.size());
    for (const BlockSpec &b : bl) {
      m.blocks.push_back(b);
      m.block_deg.push_back(deg);
      N += b.n;
    }
    m.last_b.push_back((int)m.blocks.size() - 1);
    m.last_s.push_back(N);
    return true;
  };
  if (!generation(gens[0], 1) || !generation(mid[1], narm)) return false;
  for (int a = 0; a < narm; a++) {
    m.blocks.push_back({(int)(bp - mons.begin()), 1});
    m.block_deg.push_back(1);
    N++;
    if (!generation(mid[2], 1)) return false;
  }
  if (!generation(gens[2], 1)) return false;
  m.comb = true;
  return true;
}

bool parse_branched(const std::string &c, const std::vector<std::string> &mons, MoleculeSpec &m,
                    std::string &err) {
  if (c.empty() || c[0] != '(') {
    err = "composition '" + c + "' must start with '('";
    return false;
  }
  if (c.back() == ']') {
    err = "composition '" + c + "' can not end with ']'";
    return false;
  }
  BranchParser p{c, mons, m, err};
  if (!square_positions(c, p.open, p.close)) {
    err = "composition '" + c + "': the square brackets are not balanced";
    return false;
  }
  m.blocks.clear();
  m.gnr.clear();
  m.first_s.clear();
  m.last_s.clear();
  m.first_b.clear();
  m.last_b.clear();
  p.add_generation();
  int pos = 0;
  if (!p.tree(0, pos)) return false;
  for (size_t g = 0; g < m.first_b.size(); g++)
    if (m.first_b[g] < 0) {
      err = "composition '" + c + "': an empty side group";
      return false;
    }
  // The reference then turns the molecule around (Molecule::Decomposition
  // src/molecule.cpp:1742-1767): blocks in reverse order, contour indices
  // mirrored, so first_s / last_s become the first and the last segment
  // (inclusive) of each generation, counted from the far end of the main
  // chain, and first_b / last_b its first and last block.  The propagation
  // then runs from the main chain's last segment, and the block after each
  // side group is the branch point (one segment long).
  const int CL = m.last_s[0], nb = (int)m.blocks.size();
  for (size_t g = 0; g < m.first_s.size(); g++) {
    const int fs = CL - m.first_s[g] - 1, ls = CL - m.last_s[g] - 1;
    m.first_s[g] = ls + 1;
    m.last_s[g] = fs;
    const int fb = nb - m.first_b[g] - 1, lb = nb - m.last_b[g] - 1;
    m.first_b[g] = lb;
    m.last_b[g] = fb;
  }
  std::reverse(m.blocks.begin(), m.blocks.end());
  std::reverse(m.gnr.begin(), m.gnr.end());
  m.branched = true;
  return true;
}

}  // namespace

int count_starts(const std::string &text) {
  auto lines = clean_lines(text);
  int n = 0;
  for (auto &l : lines)
    if (l.compare(0, 5, "start") == 0) n++;
  return n;
}

std::string insert_before_start(const std::string &text, int start, const std::string &line) {
  std::istringstream is(text);
  std::string raw, out;
  int n = 0;
  bool done = false;
  while (std::getline(is, raw)) {
    std::string t;
    for (char ch : raw)
      if (ch != ' ' && ch != '\t' && ch != '\r') t.push_back(ch);
    if (!done && t.compare(0, 5, "start") == 0 && ++n == start) {
      out += line + "\n";
      done = true;
    }
    out += raw + "\n";
  }
  if (!done) out += line + "\nstart\n";
  return out;
}

double scan_value(const VarScan &v, int k) {
  if (v.exponential) {
    const double f = 1.0 * k / v.num_of_steps;
    const double sgn = v.start_value < 0 ? -1.0 : 1.0;
    return sgn * std::pow(10.0, (1 - f) * std::log10(sgn * v.start_value) + f * std::log10(sgn * v.end_value));
  }
  return v.start_value + k * v.step;
}

bool parse_input(const std::string &text, int start, SystemSpec &s, std::string &err) {
  Parsed p;
  if (start < 1) start = 1;
  if (!collect(text, start, p, err)) return false;
  s = SystemSpec();
  s.sys_name = p.sys_name;
  if (!p.newton_name.empty()) s.newton_name = p.newton_name;
  std::string v;
  // ---------------- lattice (src/lattice.cpp:267-571) ----------------
  LatticeSpec &l = s.lat;
  const std::string &ln = p.lat_name;
  if (ln.empty()) {
    err = "no 'lat' block";
    return false;
  }
  if (get(p, "lat", ln, "gradients", v) && !to_int(v, l.gradients)) l.gradients = 1;
  if (l.gradients < 1 || l.gradients > 3) l.gradients = 1;
  // lattice refinement (src/lattice.cpp:296-329): FJC_choices = 3 + 2i or
  // b/l = fjc, FJC = 2 fjc + 1; inconsistent pairs keep b/l when it is the
  // smaller, else FJC_choices
  l.fjc = 1;
  l.FJC = 3;
  const bool has_fjc = get(p, "lat", ln, "FJC_choices", v);
  if (has_fjc) {
    if (!to_int(v, l.FJC) || l.FJC < 3 || (l.FJC - 3) % 2 != 0) {
      err = "FJC_choices can adopt only the values 3 + 2i, i = 0, 1, 2, ...";
      return false;
    }
    l.fjc = (l.FJC - 1) / 2;
  }
  if (get(p, "lat", ln, "b/l", v)) {
    int b;
    if (!to_int(v, b) || b < 1) {
      err = "b/l must be a positive integer";
      return false;
    }
    if (has_fjc && b != l.fjc && !(b < l.fjc)) b = l.fjc;
    l.fjc = b;
    l.FJC = 2 * b + 1;
  }
  if (l.fjc > kMaxFjcSpec) {
    err = "lattice refinement above FJC_choices 9 is not supported";
    return false;
  }
  // src/lattice.cpp:551-552 (3-gradient refinement is not restated here)
  if (l.fjc > 3 && l.gradients == 2) {
    err = "with 2 gradients FJC_choices are limited to 7";
    return false;
  }
  if (l.fjc > 2 && l.gradients == 3) {
    err = "with 3 gradients FJC_choices are limited to 5";
    return false;
  }
  if (!get(p, "lat", ln, "lattice_type", v)) {
    err = "'lattice_type' is required";
    return false;
  }
  s.lattice_type_name = v;
  {
    std::string bl;
    double b = 0;
    if (get(p, "lat", ln, "bondlength", bl) && to_real(bl, b)) l.bond_length = b;
  }
  if (v == "simple_cubic") {
    l.lattice_type = LAT_SIMPLE_CUBIC;
    l.lambda = 1.0 / 6.0;
  } else if (v == "hexagonal") {
    l.lattice_type = LAT_HEXAGONAL;
    l.lambda = 1.0 / 4.0;
  } else {
    err = "lattice_type '" + v + "' not recognised";
    return false;
  }
  // src/lattice.cpp:550
  if (l.fjc > 1 && l.lattice_type != LAT_HEXAGONAL) {
    err = "for FJC_choices > 3 the lattice_type must be hexagonal";
    return false;
  }
  l.geometry = GEOM_PLANAR;
  if (get(p, "lat", ln, "geometry", v)) {
    if (v == "planar" || v == "flat")
      l.geometry = GEOM_PLANAR;
    else if (v == "cylindrical")
      l.geometry = GEOM_CYLINDRICAL;
    else if (v == "spherical")
      l.geometry = GEOM_SPHERICAL;
    else {
      err = "geometry '" + v + "' not recognised";
      return false;
    }
  }
  if (l.gradients == 3 && l.geometry != GEOM_PLANAR) {
    err = "3 gradients require planar geometry";
    return false;
  }
  if (l.gradients == 2 && l.geometry == GEOM_SPHERICAL) {
    err = "2 gradients do not support spherical geometry";
    return false;
  }
  if (l.geometry != GEOM_PLANAR && get(p, "lat", ln, "offset_first_layer", v)) {
    if (!to_real(v, l.offset_first_layer) || l.offset_first_layer < 0) l.offset_first_layer = 0;
  }
  auto need_int = [&](const char *k, int &dst) {
    if (!get(p, "lat", ln, k, v) || !to_int(v, dst) || dst < 1) {
      err = std::string("'") + k + "' is required and must be positive";
      return false;
    }
    return true;
  };
  bool ok;
  if (l.gradients == 1) {
    if (!need_int("n_layers", l.MX)) return false;
    l.MX *= l.fjc;
    l.offset_first_layer *= l.fjc;
    l.bc[0] = get(p, "lat", ln, "lowerbound", v) ? bc_of(v, ok) : BC_MIRROR;
    if (get(p, "lat", ln, "lowerbound", v) && !ok) {
      err = "lowerbound '" + v + "' not recognised";
      return false;
    }
    l.bc[3] = get(p, "lat", ln, "upperbound", v) ? bc_of(v, ok) : BC_MIRROR;
    if (get(p, "lat", ln, "upperbound", v) && !ok) {
      err = "upperbound '" + v + "' not recognised";
      return false;
    }
  } else {
    if (!need_int("n_layers_x", l.MX) || !need_int("n_layers_y", l.MY)) return false;
    if (l.gradients == 3 && !need_int("n_layers_z", l.MZ)) return false;
    l.MX *= l.fjc;
    l.MY *= l.fjc;
    if (l.gradients == 3) l.MZ *= l.fjc;
    static const char *lo[3] = {"lowerbound_x", "lowerbound_y", "lowerbound_z"};
    static const char *hi[3] = {"upperbound_x", "upperbound_y", "upperbound_z"};
    for (int d = 0; d < l.gradients; d++) {
      for (int side = 0; side < 2; side++) {
        const char *k = side ? hi[d] : lo[d];
        int &b = l.bc[d + 3 * side];
        b = BC_MIRROR;
        if (get(p, "lat", ln, k, v)) {
          b = bc_of(v, ok);
          if (!ok) {
            err = std::string(k) + " '" + v + "' not recognised";
            return false;
          }
          if (l.gradients == 3 && b == BC_SURFACE) {
            err = "surface boundaries are not available with 3 gradients";
            return false;
          }
          if (l.gradients == 2 && l.geometry != GEOM_PLANAR && b == BC_PERIODIC) {
            err = "periodic boundaries need planar geometry in 2 gradients";
            return false;
          }
        }
      }
      if ((l.bc[d] == BC_PERIODIC) != (l.bc[d + 3] == BC_PERIODIC)) {
        err = "a periodic boundary must be set on both sides";
        return false;
      }
    }
  }
  if (l.gradients == 1 && ((l.bc[0] == BC_PERIODIC) != (l.bc[3] == BC_PERIODIC))) {
    err = "a periodic boundary must be set on both sides";
    return false;
  }
  l.stencil_full = true;
  if (get(p, "lat", ln, "stencil_full", v)) {
    bool b;
    if (to_bool(v, b)) l.stencil_full = b;
  }
  if (l.fjc > 1 && l.gradients == 3 && !l.stencil_full) {
    err = "refined 3-gradient lattices need stencil_full";
    return false;
  }
  // PutM (src/LGrad1.cpp:180-190, src/LGrad2.cpp:101-111, src/LGrad3.cpp:15-23)
  if (l.gradients == 1) {
    l.JX = 1;
    l.JY = 0;
    l.JZ = 0;
    const int f = l.fjc;
    l.M = l.MX + 2 * f;
    if (l.geometry == GEOM_PLANAR) l.volume = l.MX / f;
    if (l.geometry == GEOM_SPHERICAL)
      l.volume = 4.0 / 3.0 * kPIE * (std::pow(l.MX + l.offset_first_layer, 3) - std::pow(l.offset_first_layer, 3)) /
                 f / f / f;
    if (l.geometry == GEOM_CYLINDRICAL)
      l.volume = kPIE * (std::pow(l.MX + l.offset_first_layer, 2) - std::pow(l.offset_first_layer, 2)) / f / f;
  } else if (l.gradients == 2) {
    l.JX = l.MY + 2 * l.fjc;
    l.JY = 1;
    l.JZ = 0;
    l.M = (l.MX + 2 * l.fjc) * (l.MY + 2 * l.fjc);
    if (l.geometry == GEOM_CYLINDRICAL)
      l.volume = l.MY * kPIE * (std::pow(l.MX + l.offset_first_layer, 2) - std::pow(l.offset_first_layer, 2));
    else
      l.volume = (double)l.MX * l.MY;
  } else {
    const int f = l.fjc;
    l.JX = (l.MZ + 2 * f) * (l.MY + 2 * f);
    l.JY = l.MZ + 2 * f;
    l.JZ = 1;
    l.M = (l.MX + 2 * f) * (l.MY + 2 * f) * (l.MZ + 2 * f);
    l.volume = (double)l.MX * l.MY * l.MZ;
  }
  boundary_indices(l);

  // ---------------- segments ----------------
  bool frozen_at_bound[6] = {false, false, false, false, false, false};
  const int nseg = (int)p.mons.size();
  if (nseg == 0) {
    err = "no 'mon' defined";
    return false;
  }
  s.seg.resize(nseg);
  for (int i = 0; i < nseg; i++) {
    SegmentSpec &sg = s.seg[i];
    sg.name = p.mons[i];
    sg.mask.assign(l.M, 0);
    if (!get(p, "mon", sg.name, "freedom", v)) {
      err = "mon '" + sg.name + "' has no 'freedom'";
      return false;
    }
    s.seg_freedom_name.push_back(v);
    {
      // valence and epsilon keep an out-of-range value, as the reference
      // does after its warning (Segment::CheckInput src/segment.cpp:1056-1067)
      std::string vv;
      double val = 0;
      sg.valence = get(p, "mon", sg.name, "valence", vv) && to_real(vv, val) ? val : 0.0;
      s.seg_valence.push_back(sg.valence);
      sg.epsilon = 80.0;
      if (get(p, "mon", sg.name, "epsilon", vv) && !to_real(vv, sg.epsilon)) sg.epsilon = 80.0;
      if (sg.valence != 0) {
        const double b = l.bond_length / l.fjc;
        if (b < 1e-12 || b > 1e-8) {
          err = "mon '" + sg.name + "' is charged: set lat bondlength in 1e-12..1e-8 m";
          return false;
        }
      }
      if (get(p, "mon", sg.name, "e.psi0/kT", vv)) {
        sg.fixed_psi0 = true;
        if (!to_real(vv, sg.psi0)) sg.psi0 = 0.0;
        if (sg.psi0 != 0 && sg.valence != 0) {
          err = "mon '" + sg.name + "': set only one of 'valence' and 'e.psi0/kT'";
          return false;
        }
        if (sg.psi0 != 0 && v != "frozen") {
          err = "mon '" + sg.name + "': a potential 'e.psi0/kT' needs freedom 'frozen'";
          return false;
        }
        if (sg.psi0 < -25 || sg.psi0 > 25) {
          err = "mon '" + sg.name + "': 'e.psi0/kT' out of range -25..25";
          return false;
        }
      }
    }
    if (v == "free")
      sg.freedom = SEG_FREE;
    else if (v == "frozen")
      sg.freedom = SEG_FROZEN;
    else if (v == "pinned")
      sg.freedom = SEG_PINNED;
    else if (v == "tagged")
      sg.freedom = SEG_TAGGED;
    else {
      err = "mon '" + sg.name + "': freedom '" + v + "' is not supported (free, frozen, pinned, tagged)";
      return false;
    }
    if (sg.freedom == SEG_TAGGED) {
      if (!get(p, "mon", sg.name, "tagged_range", v)) {
        err = "mon '" + sg.name + "': tagged segments need 'tagged_range'";
        return false;
      }
      if (!parse_range(l, v, sg.mask, err)) return false;
    }
    if (sg.freedom == SEG_FROZEN) {
      if (!get(p, "mon", sg.name, "frozen_range", v)) {
        err = "mon '" + sg.name + "': frozen segments need 'frozen_range'";
        return false;
      }
      if (!parse_range(l, v, sg.mask, err)) return false;
      // a frozen layer given by a boundary keyword backs that surface
      // (Segment::ParseFreedoms src/segment.cpp:418-440)
      static const char *axes[4] = {"", "_x", "_y", "_z"};
      for (int a = 0; a < 4; a++) {
        if (v == std::string("lowerbound") + axes[a]) frozen_at_bound[a == 0 ? 0 : a - 1] = true;
        if (v == std::string("upperbound") + axes[a]) frozen_at_bound[a == 0 ? 3 : a + 2] = true;
        if (sg.valence != 0 && (v == std::string("lowerbound") + axes[a] || v == std::string("upperbound") + axes[a])) {
          // System::CheckInput src/system.cpp:1193-1200
          err = "mon '" + sg.name + "': a charged frozen segment may not sit in a boundary";
          return false;
        }
      }
    }
    if (sg.freedom == SEG_PINNED) {
      if (!get(p, "mon", sg.name, "pinned_range", v)) {
        err = "mon '" + sg.name + "': pinned segments need 'pinned_range'";
        return false;
      }
      if (!parse_range(l, v, sg.mask, err)) return false;
    }
  }
  // a surface boundary needs a frozen segment on it (System::CheckInput
  // src/system.cpp:1193-1209)
  static const char *side_name[6] = {"lowerboundary in x", "lowerboundary in y", "lowerboundary in z",
                                     "upperboundary in x", "upperboundary in y", "upperboundary in z"};
  for (int b = 0; b < 6; b++)
    if (l.bc[b] == BC_SURFACE && !frozen_at_bound[b]) {
      err = std::string("lonely 'surface': specify a segment with frozen_range including the ") + side_name[b];
      return false;
    }
  // chi matrix (Segment::CheckInput src/segment.cpp:1105-1119,
  // System::CheckChi_values src/system.cpp:1718-1800): unset entries take
  // the transposed value, otherwise 0.
  const double unset = -999.0;
  std::vector<double> chi(nseg * nseg, unset);
  for (int i = 0; i < nseg; i++)
    for (int k = 0; k < nseg; k++)
      if (get(p, "mon", s.seg[i].name, "chi_" + s.seg[k].name, v)) {
        double c;
        if (!to_real(v, c)) {
          err = "chi value '" + v + "' is not a number";
          return false;
        }
        chi[i * nseg + k] = (i == k) ? 0.0 : c;
      }
  for (int i = 0; i < nseg; i++)
    for (int k = 0; k < nseg; k++) {
      double &a = chi[i * nseg + k], &b = chi[k * nseg + i];
      if (a == unset && b == unset) {
        a = b = 0.0;
      } else if (a != unset && b == unset) {
        b = a;
      }
    }
  for (int i = 0; i < nseg; i++)
    for (int k = i + 1; k < nseg; k++)
      if (chi[i * nseg + k] != chi[k * nseg + i]) {
        err = "chi(" + s.seg[i].name + "," + s.seg[k].name + ") is not symmetric";
        return false;
      }
  for (int i = 0; i < nseg; i++) s.seg[i].chi.assign(chi.begin() + i * nseg, chi.begin() + (i + 1) * nseg);

  // ---------------- molecules ----------------
  const int nmol = (int)p.mols.size();
  if (nmol == 0) {
    err = "no 'mol' defined";
    return false;
  }
  s.mol.resize(nmol);
  bool any_markov2 = false;
  for (int i = 0; i < nmol; i++) {
    MoleculeSpec &m = s.mol[i];
    m.name = p.mols[i];
    if (!get(p, "mol", m.name, "composition", v)) {
      err = "mol '" + m.name + "' has no composition";
      return false;
    }
    s.mol_composition.push_back(v);
    if (!v.empty() && v[0] == '@') {
      if (v.compare(0, 6, "@comb(") == 0) {
        if (!parse_comb(v, p.mons, m, err)) return false;
      } else if (v.compare(0, 6, "@dend(") == 0) {
        if (!parse_dendrimer(v, p.mons, m, err)) return false;
      } else {
        err = "composition '" + v + "': of the @ architectures only @dend (symmetric dendrimers) and @comb are "
              "supported";
        return false;
      }
    } else if (v.find('[') != std::string::npos) {
      if (!parse_branched(v, p.mons, m, err)) return false;
    } else if (!parse_composition(v, p.mons, m.blocks, err)) {
      return false;
    }
    if (m.block_deg.size() != m.blocks.size()) m.block_deg.assign(m.blocks.size(), 1);
    m.chainlength = 0;
    for (size_t bi = 0; bi < m.blocks.size(); bi++) {
      const BlockSpec &b = m.blocks[bi];
      m.chainlength += b.n * m.block_deg[bi];
      if (s.seg[b.seg].freedom == SEG_FROZEN) {
        err = "mol '" + m.name + "' contains a frozen segment";
        return false;
      }
      if (s.seg[b.seg].freedom == SEG_PINNED) m.pinned = true;
      if (s.seg[b.seg].freedom == SEG_TAGGED) {
        if (m.tagged) {
          err = "mol '" + m.name + "': one tagged segment per molecule is supported";
          return false;
        }
        m.tagged = true;
        if (b.n != 1) {
          err = "mol '" + m.name + "': a tagged block must be a single segment";
          return false;
        }
      }
      auto it = std::find(m.molmon.begin(), m.molmon.end(), b.seg);
      if (it == m.molmon.end()) {
        m.block_molmon.push_back((int)m.molmon.size());
        m.molmon.push_back(b.seg);
      } else {
        m.block_molmon.push_back((int)(it - m.molmon.begin()));
      }
    }
    // ring : true (Molecule::CheckInput src/molecule.cpp:584-625): the first
    // and last blocks of a linear chain, one segment each of the same type,
    // are merged and the chain is one segment shorter; a chain that does not
    // qualify stays linear, as in the reference
    if (get(p, "mol", m.name, "ring", v)) {
      bool r = false;
      if (!to_bool(v, r)) {
        err = "mol '" + m.name + "': ring is either 'true' or 'false'";
        return false;
      }
      if (r && m.scheduled()) {
        err = "mol '" + m.name + "': rings of branched molecules, dendrimers and combs are not supported";
        return false;
      }
      const size_t nb = m.blocks.size();
      if (r && nb >= 2 && m.blocks[0].seg == m.blocks[nb - 1].seg && m.blocks[0].n == 1 && m.blocks[nb - 1].n == 1) {
        m.ring = true;
        m.chainlength--;
      }
    }
    if (m.tagged) {
      // n is the number of tagged sites; freedom, theta, n, phibulk are
      // ignored (src/system.cpp:606-611, src/molecule.cpp:578-580)
      if (m.pinned || m.scheduled()) {
        err = "mol '" + m.name + "': tagged molecules must be linear and unpinned";
        return false;
      }
      if (m.chainlength < 2) {  // F and mu divide by N - 1
        err = "mol '" + m.name + "': a tagged molecule needs segments besides the tagged one";
        return false;
      }
      int npos = 0;
      for (const BlockSpec &b : m.blocks)
        if (s.seg[b.seg].freedom == SEG_TAGGED)
          for (int z = 0; z < l.M; z++) npos += s.seg[b.seg].mask[z];
      m.freedom = MOL_TAGGED;
      m.n = npos;
      m.theta = m.n * m.chainlength;
      s.mol_freedom_name.push_back("tagged");
    }
    if (!m.tagged) {
      if (!get(p, "mol", m.name, "freedom", v)) {
        err = "mol '" + m.name + "' has no freedom";
        return false;
      }
      s.mol_freedom_name.push_back(v);
      if (v == "free")
        m.freedom = MOL_FREE;
      else if (v == "restricted")
        m.freedom = MOL_RESTRICTED;
      else if (v == "solvent")
        m.freedom = MOL_SOLVENT;
      else if (v == "neutralizer")
        m.freedom = MOL_NEUTRALIZER;
      else {
        err = "mol '" + m.name + "': freedom '" + v + "' is not supported (free, restricted, solvent, neutralizer)";
        return false;
      }
      if (m.pinned && m.freedom != MOL_RESTRICTED) {
        err = "mol '" + m.name + "' is pinned and must be restricted";
        return false;
      }
      if (m.freedom == MOL_FREE) {
        if (!get(p, "mol", m.name, "phibulk", v) || !to_real(v, m.phibulk) || m.phibulk < 0 || m.phibulk > 1) {
          err = "mol '" + m.name + "': free molecules need phibulk in 0..1";
          return false;
        }
      }
      if (m.freedom == MOL_RESTRICTED) {
        bool hn = get(p, "mol", m.name, "n", v);
        std::string vt;
        bool ht = get(p, "mol", m.name, "theta", vt);
        if (hn == ht) {
          err = "mol '" + m.name + "': restricted molecules need exactly one of 'n' or 'theta'";
          return false;
        }
        // a ring's n and theta convert with the composition's length: the
        // reference reads them before it shortens the chain
        // (Molecule::CheckInput src/molecule.cpp:545-546 before :609)
        const int nin = m.chainlength + (m.ring ? 1 : 0);
        if (hn) {
          if (!to_real(v, m.n)) {
            err = "mol '" + m.name + "': bad n";
            return false;
          }
          m.theta = m.n * nin;
        } else {
          if (!to_real(vt, m.theta)) {
            err = "mol '" + m.name + "': bad theta";
            return false;
          }
          m.n = m.theta / nin;
        }
        if (m.theta < 0 || m.theta > l.volume) {
          err = "mol '" + m.name + "': theta out of range 0..volume";
          return false;
        }
      }
    }
    m.markov = 1;
    if (get(p, "mol", m.name, "Markov", v) && (!to_int(v, m.markov) || m.markov < 1 || m.markov > 2)) {
      err = "mol '" + m.name + "': Markov must be 1 or 2";
      return false;
    }
    m.k_stiff = 0.0;
    if (get(p, "mol", m.name, "k_stiff", v) && (!to_real(v, m.k_stiff) || m.k_stiff < 0 || m.k_stiff > 10)) {
      err = "mol '" + m.name + "': k_stiff out of range 0..10";
      return false;
    }
    m.save_memory = false;
    if (get(p, "mol", m.name, "save_memory", v)) to_bool(v, m.save_memory);
    if (m.ring && (m.markov != 1 || m.save_memory || m.pinned || m.tagged || l.fjc != 1)) {
      err = "mol '" + m.name +
            "': rings are supported for first-order chains on lattices with fjc 1, without save_memory, "
            "pinned or tagged segments";
      return false;
    }
    m.size = 1;
    if (m.markov == 2) {
      const bool sc1 = l.lattice_type == LAT_SIMPLE_CUBIC && l.fjc == 1;
      // hexagonal: planar with any refinement, curved with fjc = 1
      const bool hexp = l.lattice_type == LAT_HEXAGONAL && (l.geometry == GEOM_PLANAR || l.fjc == 1);
      const bool ok1 = l.gradients == 1 && (sc1 || hexp);
      const bool ok23 = l.gradients >= 2 && sc1;
      if (!(ok1 || ok23)) {
        err = "mol '" + m.name + "': Markov 2 is supported on simple-cubic lattices without refinement, on "
              "planar 1-gradient hexagonal lattices and on curved ones without refinement";
        return false;
      }
      any_markov2 = true;
      // direction components (Molecule::CheckInput src/molecule.cpp:648-670):
      // FJC in 1 gradient, 2 FJC - 1 in 2 and 6 in 3 gradients, where Markov 2
      // also turns stencil_full off for the whole lattice
      m.size = l.gradients == 1 ? l.FJC : (l.gradients == 2 ? 2 * l.FJC - 1 : 6);
      if (l.gradients >= 2) l.stencil_full = false;
      // Bending weights (Molecule::AllocateMemory src/molecule.cpp:133-153)
      m.P.assign(l.FJC, 0.0);
      double Q = 0;
      for (int k = 0; k < l.FJC - 1; k++) {
        double a = k * kPIE / (l.FJC - 1);
        m.P[k] = std::exp(-0.5 * m.k_stiff * a * a);
        if (k > 0)
          Q += (l.lattice_type == LAT_HEXAGONAL ? 2 : 4) * m.P[k];
        else
          Q = m.P[k];
      }
      m.P[l.FJC - 1] = 0;
      if (l.lattice_type == LAT_HEXAGONAL && !l.stencil_full) Q *= 2.0;
      for (int k = 0; k < l.FJC; k++) m.P[k] /= Q;
    }
    if (m.freedom == MOL_SOLVENT) {
      if (s.solvent >= 0) {
        err = "more than one solvent molecule";
        return false;
      }
      s.solvent = i;
    }
  }
  if (s.solvent < 0) {
    if (!(nmol == 1 && s.mol[0].pinned)) {
      err = "no solvent molecule (freedom : solvent) found";
      return false;
    }
  }
  // Electrostatics (System::CheckInput src/system.cpp:701-729,
  // Molecule::Charge / IsCharged src/molecule.cpp:1866-1898): a molecule is
  // charged when its valences do not sum to zero; one charged molecule turns
  // the potential on for the whole system.
  s.mol_charge.assign(nmol, 0.0);
  s.charged = false;
  s.neutralizer = -1;
  for (int i = 0; i < nmol; i++) {
    const MoleculeSpec &m = s.mol[i];
    double q = 0;
    for (const BlockSpec &b : m.blocks) q += s.seg[b.seg].valence * b.n;
    s.mol_charge[i] = q / m.chainlength;
    if (q != 0 && m.ring) {
      err = "mol '" + m.name + "': charged rings are not supported";
      return false;
    }
    if (q != 0) {
      if (m.dendrimer || m.comb) {
        err = "mol '" + m.name + "': charged dendrimers and combs are not supported";
        return false;
      }
      s.charged = true;
    }
    if (m.freedom == MOL_NEUTRALIZER) {
      if (s.neutralizer >= 0) {
        err = "more than one neutralizer molecule";
        return false;
      }
      s.neutralizer = i;
    }
  }
  if (s.neutralizer >= 0 && !s.charged) {
    err = "mol '" + s.mol[s.neutralizer].name + "': a neutralizer needs charged molecules";
    return false;
  }
  if (s.charged) {
    if (l.gradients != 1 && l.fjc != 1) {
      err = "charged systems on 2- and 3-gradient lattices need fjc = 1";
      return false;
    }
    if (s.neutralizer < 0) {
      double qb = 0;
      bool needed = false;
      for (int i = 0; i < nmol; i++) {
        const MoleculeSpec &m = s.mol[i];
        if (m.freedom == MOL_FREE) qb += s.mol_charge[i] * m.phibulk;
        if (m.freedom == MOL_RESTRICTED && s.mol_charge[i] != 0 && !m.pinned) needed = true;
      }
      if (needed || std::fabs(qb) > 1e-4) {
        err = "a neutralizer is needed: un-pinned restricted molecules are charged or the free molecules' bulk "
              "is not neutral";
        return false;
      }
    } else if (s.solvent >= 0 && s.mol_charge[s.neutralizer] == s.mol_charge[s.solvent]) {
      err = "the neutralizer carries the solvent's charge";
      return false;
    }
    for (int i = 0; i < nseg; i++) {
      if (s.seg[i].epsilon != s.seg[0].epsilon) s.grad_epsilon = true;
      if (s.seg[i].freedom == SEG_FROZEN && s.seg[i].fixed_psi0) s.fixed_psi0 = true;
    }
  }
  compute_lambdas(l, any_markov2);

  // KSAM (System::generate_mask src/system.cpp:204-273)
  s.ksam.assign(l.M, 1);
  for (int i = 0; i < nseg; i++)
    if (s.seg[i].freedom == SEG_FROZEN || s.seg[i].freedom == SEG_TAGGED)
      for (int z = 0; z < l.M; z++)
        if (s.seg[i].mask[z]) s.ksam[z] = 0;

  // Iteration list (System::MakeItsLists src/system.cpp:429-500 with
  // IsUnique src/system.cpp: segments with identical chi rows share a field).
  std::vector<int> sysmon;
  for (auto &m : s.mol)
    for (int sg : m.molmon) {
      if (s.seg[sg].freedom == SEG_TAGGED) continue;  // no field (SysTagList, src/system.cpp:466-474)
      if (std::find(sysmon.begin(), sysmon.end(), sg) != sysmon.end()) continue;
      sysmon.push_back(sg);
      int copy = -1;
      for (size_t k = 0; k < s.itmon.size() && copy < 0; k++)
        if (s.seg[s.itmon[k]].chi == s.seg[sg].chi) copy = (int)k;
      if (copy < 0) {
        s.seg[sg].field_block = (int)s.itmon.size();
        s.itmon.push_back(sg);
      } else {
        s.seg[sg].field_block = copy;
      }
    }

  // ---------------- solver (src/solve_scf.cpp:114-281) ----------------
  const std::string &nn = p.newton_name;
  SolverSpec &so = s.solver;
  if (!nn.empty()) {
    if (get(p, "newton", nn, "iterationlimit", v)) {
      if (!to_int(v, so.iterationlimit) || so.iterationlimit < 0 || so.iterationlimit > 1000000)
        so.iterationlimit = 1000;
    }
    if (get(p, "newton", nn, "deltamax", v)) {
      if (!to_real(v, so.deltamax) || so.deltamax < 0 || so.deltamax > 100) so.deltamax = 0.1;
    }
    if (get(p, "newton", nn, "tolerance", v)) {
      if (!to_real(v, so.tolerance)) so.tolerance = 1e-7;
      if (so.tolerance < 1e-16 || so.tolerance > 10) so.tolerance = 1e-5;
    }
    if (get(p, "newton", nn, "method", v)) {
      s.method_name = v;
      if (v == "DIIS")
        so.method = SOLVER_DIIS;
      else if (v == "pseudohessian")
        so.method = SOLVER_PSEUDOHESSIAN;
      else if (v == "hessian")
        so.method = SOLVER_HESSIAN;
      else if (v == "LBFGS")
        so.method = SOLVER_LBFGS;
      else if (v == "BRR")
        so.method = SOLVER_BRR;
      else {
        err = "newton method '" + v + "' is not supported (DIIS, pseudohessian, hessian, LBFGS, BRR)";
        return false;
      }
    }
    // Solve_scf::CheckInput (src/solve_scf.cpp:170-175,246-250): LBFGS keeps
    // m = 6 correction pairs unless given (0..1000), the other methods a DIIS
    // history of 10 (0..100); deltamin defaults to 0, out of 0..100 it is
    // deltamax / 1e5
    if (so.method == SOLVER_LBFGS) {
      so.m = 6;
      if (get(p, "newton", nn, "m", v))
        if (!to_int(v, so.m) || so.m < 0 || so.m > 1000) so.m = 6;
    } else if (so.method == SOLVER_BRR) {
      // BRR keeps m = 10 rank-one terms, 0..1000 (src/solve_scf.cpp:251-255)
      so.m = 10;
      if (get(p, "newton", nn, "m", v))
        if (!to_int(v, so.m) || so.m < 0 || so.m > 1000) so.m = 10;
    } else if (get(p, "newton", nn, "m", v)) {
      if (!to_int(v, so.m) || so.m < 0 || so.m > 100) so.m = 10;
    }
    if (get(p, "newton", nn, "deltamin", v)) {
      if (!to_real(v, so.deltamin) || so.deltamin < 0 || so.deltamin > 100) so.deltamin = so.deltamax / 100000;
    }
    if (get(p, "newton", nn, "stop_criterion", v)) {
      so.max_g = (v == "max_of_element_of_|g|");
      if (so.max_g) s.stop_criterion_name = v;
    }
  }
  if (get(p, "sys", p.sys_name, "initial_guess", v)) s.initial_guess = v;
  if (get(p, "sys", p.sys_name, "guess_inputfile", v)) s.guess_inputfile = v;
  if (get(p, "sys", p.sys_name, "final_guess", v) && (v == "next_problem" || v == "file")) s.final_guess = v;
  if (get(p, "sys", p.sys_name, "guess_outputfile", v)) s.guess_outputfile = v;
  s.start = start;
  // var scan (Variate::CheckInput, src/variate.cpp:64-345; the scan counts
  // of Molecule::PutVarScan / Segment::PutVarScan, src/molecule.cpp:785-816,
  // src/segment.cpp:1402-1438, truncated to an integer as there)
  for (auto &ln : p.lines) {
    if (ln.key != "var" || ln.prop != "scan" || s.scan.present) continue;
    VarScan &vs = s.scan;
    const size_t dash = ln.name.find('-');
    if (dash == std::string::npos) {
      err = "var '" + ln.name + "': expected <kind>-<name>";
      return false;
    }
    vs.kind = ln.name.substr(0, dash);
    vs.name = ln.name.substr(dash + 1);
    vs.prop = ln.value;
    bool found = false;
    if (vs.kind == "mol") {
      for (auto &m : s.mol)
        if (m.name == vs.name && (vs.prop == "theta" || vs.prop == "n" || vs.prop == "phibulk")) {
          found = true;
          vs.start_value = vs.prop == "theta" ? m.theta : (vs.prop == "n" ? m.n : m.phibulk);
        }
    } else if (vs.kind == "mon" && (vs.prop == "valence" || vs.prop == "ePsi0/kT")) {
      // Segment::PutVarInfo src/segment.cpp:1367-1369
      for (size_t i = 0; i < s.seg.size(); i++)
        if (s.seg[i].name == vs.name) {
          found = true;
          vs.start_value = vs.prop == "valence" ? s.seg[i].valence : s.seg[i].psi0;
        }
    } else if (vs.kind == "mon" && vs.prop.compare(0, 4, "chi_") == 0) {
      const std::string other = vs.prop.substr(4);
      for (size_t i = 0; i < s.seg.size(); i++)
        if (s.seg[i].name == vs.name)
          for (size_t j = 0; j < s.seg.size(); j++)
            if (s.seg[j].name == other) {
              found = true;
              vs.start_value = s.seg[i].chi[j];
            }
    }
    if (!found) {
      err = "var scan of " + ln.name + " : " + vs.prop +
            " is not supported (mol theta, n, phibulk; mon chi_X, valence, ePsi0/kT)";
      return false;
    }
    std::string sv;
    if (!get(p, "var", ln.name, "end_value", sv) || !to_real(sv, vs.end_value)) {
      err = "var scan needs an end_value";
      return false;
    }
    vs.exponential = get(p, "var", ln.name, "scale", sv) && sv == "exponential";
    if (vs.exponential) {
      if (vs.kind == "mon") {
        err = "a chi scan is linear only";
        return false;
      }
      if (get(p, "var", ln.name, "steps", sv) && !to_int(sv, vs.steps)) vs.steps = 1;
      if (vs.steps < 1 || vs.end_value * vs.start_value <= 0) {
        err = "exponential var scan needs steps >= 1 and end_value, start value of one sign";
        return false;
      }
      const double r = vs.end_value > vs.start_value ? vs.end_value / vs.start_value : vs.start_value / vs.end_value;
      vs.num_of_steps = (int)(vs.steps * std::log10(r));
    } else {
      if (!get(p, "var", ln.name, "step", sv) || !to_real(sv, vs.step) || vs.step == 0) {
        err = "linear var scan needs a nonzero step";
        return false;
      }
      const double q = (vs.end_value - vs.start_value) / vs.step;
      if (q < 0) {
        err = "var scan: (end_value - start value) / step is negative";
        return false;
      }
      vs.num_of_steps = (int)q;
    }
    vs.present = true;
  }
  // var search with its target (see VarSearch)
  {
    VarSearch &vs = s.search;
    int n_search = 0, n_target = 0;
    for (auto &ln : p.lines) {
      if (ln.key != "var") continue;
      const size_t dash = ln.name.find('-');
      if (dash == std::string::npos) continue;
      const std::string kind = ln.name.substr(0, dash), name = ln.name.substr(dash + 1);
      if (ln.prop == "search") {
        n_search++;
        if (kind != "mol" || (ln.value != "theta" && ln.value != "n" && ln.value != "phibulk")) {
          err = "var search of " + ln.name + " : " + ln.value + " is not supported (mol theta, n, phibulk)";
          return false;
        }
        for (size_t i = 0; i < s.mol.size(); i++)
          if (s.mol[i].name == name) {
            const MoleculeSpec &m = s.mol[i];
            vs.mol = (int)i;
            vs.start_value = ln.value == "theta" ? m.theta : (ln.value == "n" ? m.n : m.phibulk);
          }
        if (vs.mol < 0) {
          err = "var search: no molecule " + name;
          return false;
        }
        vs.name = name;
        vs.prop = ln.value;
      }
      const bool sys_target = kind == "sys" && (ln.prop == "free_energy" || ln.prop == "grand_potential");
      const bool mol_target =
          kind == "mol" && (ln.prop == "theta" || ln.prop == "n" || ln.prop == "phibulk" || ln.prop == "mu");
      if (sys_target || mol_target) {
        n_target++;
        vs.target_kind = kind;
        vs.target_prop = ln.prop;
        if (!to_real(ln.value, vs.target_value)) {
          err = "var target " + ln.name + " : " + ln.prop + " needs a number";
          return false;
        }
        if (mol_target) {
          for (size_t i = 0; i < s.mol.size(); i++)
            if (s.mol[i].name == name) vs.target_mol = (int)i;
          if (vs.target_mol < 0) {
            err = "var target: no molecule " + name;
            return false;
          }
        }
      } else if (kind == "sys" && ln.prop == "Laplace_pressure") {
        err = "var target Laplace_pressure is not supported";
        return false;
      }
    }
    if (n_search > 1 || n_target > 1) {
      err = "at most one var search and one target (Variate, as in the reference)";
      return false;
    }
    if (n_search != n_target) {
      err = n_search ? "lonely search: a var search needs a target" : "lonely target: a var target needs a search";
      return false;
    }
    if (n_search) {
      vs.present = true;
      vs.super_tolerance = 10 * so.tolerance;
      vs.super_iterationlimit = so.iterationlimit / 10;
      if (get(p, "newton", nn, "super_tolerance", v)) to_real(v, vs.super_tolerance);
      if (get(p, "newton", nn, "super_iterationlimit", v)) to_int(v, vs.super_iterationlimit);
      if (get(p, "newton", nn, "super_deltamax", v)) to_real(v, vs.super_deltamax);
      if (s.scan.present) {
        err = "a var search and a var scan in one start block are not supported";
        return false;
      }
    }
  }
  // output requests (Output::CheckInput / Load, src/output.cpp:32-178):
  // `output : kal | pro | vtk | vec | pos : option : value` and
  // `kal | pro | vtk | vec : key : name : prop`
  // item lines in input order; name `*` expands to every molecule / segment
  OutputSpec &out = s.out;
  for (auto &ln : p.lines) {
    if (ln.key == "out_info" && ln.name == "folder" && ln.prop == "path" && !ln.value.empty())
      out.path = ln.value.back() == '/' ? ln.value : ln.value + "/";
    if (ln.key != "output" || (ln.name != "kal" && ln.name != "pro" && ln.name != "vtk" && ln.name != "vec" &&
                               ln.name != "pos"))
      continue;
    OutputRequest &r = ln.name == "kal"   ? out.kal
                       : ln.name == "pro" ? out.pro
                       : ln.name == "vtk" ? out.vtk
                       : ln.name == "vec" ? out.vec
                                          : out.pos;
    r.present = true;
    bool b;
    if (ln.prop == "append" && to_bool(ln.value, b)) {
      r.append = b;
      r.append_given = true;
    }
    if (ln.prop == "write_bounds" && to_bool(ln.value, b)) r.write_bounds = b;
    if (ln.prop == "write" && to_bool(ln.value, b)) r.write = b;
    if (ln.prop == "use_output_folder" && to_bool(ln.value, b)) out.use_output_folder = b;
    if (ln.prop == "header_separator")
      r.sep = ln.value == "classic" ? ":" : (ln.value.size() == 1 ? ln.value : "_");
  }
  out.pro.append = false;  // a profile file is always rewritten
  out.vtk.append = false;
  out.vec.append = true;  // a vec file is always appended to (src/output.cpp:115-124, 452-455)
  for (auto &ln : p.lines) {
    if (ln.key != "kal" && ln.key != "pro" && ln.key != "vtk" && ln.key != "vec") continue;
    OutputRequest &r = ln.key == "kal" ? out.kal : (ln.key == "pro" ? out.pro : (ln.key == "vtk" ? out.vtk : out.vec));
    if (!r.present) continue;
    const std::string &k = ln.name;
    if (ln.prop == "*" && (k == "mol" || k == "mon")) {
      if (k == "mol")
        for (auto &m : s.mol) r.items.push_back({k, m.name, ln.value});
      else
        for (auto &g : s.seg) r.items.push_back({k, g.name, ln.value});
    } else {
      r.items.push_back({k, ln.prop == "*" && k == "sys" ? s.sys_name : ln.prop, ln.value});
    }
  }
  return true;
}

std::string describe_json(const SystemSpec &s) {
  std::ostringstream o;
  auto str = [&](const std::string &v) { o << '"' << v << '"'; };
  o << "{\"sys_name\":";
  str(s.sys_name);
  o << ",\"newton_name\":";
  str(s.newton_name);
  const LatticeSpec &l = s.lat;
  o << ",\"gradients\":" << l.gradients << ",\"geometry\":" << l.geometry << ",\"fjc\":" << l.fjc << ",\"MX\":" << l.MX
    << ",\"MY\":" << l.MY << ",\"MZ\":" << l.MZ << ",\"M\":" << l.M << ",\"JX\":" << l.JX
    << ",\"JY\":" << l.JY << ",\"nvar\":" << s.nvar() << ",\"solvent\":" << s.solvent;
  o << ",\"charged\":" << (s.charged ? "true" : "false") << ",\"neutralizer\":" << s.neutralizer
    << ",\"grad_epsilon\":" << (s.grad_epsilon ? "true" : "false")
    << ",\"fixed_psi0\":" << (s.fixed_psi0 ? "true" : "false") << ",\"psi_block\":"
    << (s.charged ? (int)s.itmon.size() : -1);
  o << ",\"segments\":[";
  for (size_t i = 0; i < s.seg.size(); i++) {
    if (i) o << ",";
    o << "{\"name\":";
    str(s.seg[i].name);
    o << ",\"freedom\":" << s.seg[i].freedom << ",\"field_block\":" << s.seg[i].field_block
      << ",\"valence\":" << s.seg[i].valence << ",\"epsilon\":" << s.seg[i].epsilon << "}";
  }
  o << "],\"molecules\":[";
  for (size_t i = 0; i < s.mol.size(); i++) {
    const MoleculeSpec &m = s.mol[i];
    if (i) o << ",";
    o << "{\"name\":";
    str(m.name);
    o << ",\"chainlength\":" << m.chainlength << ",\"freedom\":" << m.freedom << ",\"markov\":" << m.markov
      << ",\"pinned\":" << (m.pinned ? "true" : "false") << ",\"ring\":" << (m.ring ? "true" : "false")
      << ",\"save_memory\":" << (m.save_memory ? "true" : "false")
      << ",\"size\":" << m.size << ",\"blocks\":" << m.blocks.size() << ",\"molmon\":" << m.molmon.size();
    if (m.scheduled()) {
      // the branched molecule's or dendrimer's blocks, generations and row schedule
      auto ints = [&](const std::vector<int> &v) {
        o << "[";
        for (size_t k = 0; k < v.size(); k++) o << (k ? "," : "") << v[k];
        o << "]";
      };
      o << ",\"branch\":{\"block_seg\":[";
      for (size_t k = 0; k < m.blocks.size(); k++) o << (k ? "," : "") << m.blocks[k].seg;
      o << "],\"block_n\":[";
      for (size_t k = 0; k < m.blocks.size(); k++) o << (k ? "," : "") << m.blocks[k].n;
      o << "],\"block_molmon\":";
      ints(m.block_molmon);
      o << ",\"dendrimer\":" << (m.dendrimer ? "true" : "false") << ",\"comb\":" << (m.comb ? "true" : "false")
        << ",\"block_deg\":";
      ints(m.block_deg);
      o << ",\"n_arm\":";
      ints(m.n_arm);
      o << ",\"gnr\":";
      ints(m.gnr);
      o << ",\"first_b\":";
      ints(m.first_b);
      o << ",\"last_b\":";
      ints(m.last_b);
      o << ",\"first_s\":";
      ints(m.first_s);
      o << ",\"last_s\":";
      ints(m.last_s);
      BranchPlan plan;
      std::string e;
      if (build_branch_plan(m, plan, e)) {
        auto ops = [&](const std::vector<BranchOp> &v) {
          o << "[";
          for (size_t k = 0; k < v.size(); k++)
            o << (k ? "," : "") << "[" << v[k].kind << "," << v[k].a << "," << v[k].b << "," << v[k].c << ","
              << v[k].seg << "," << v[k].f << "]";
          o << "]";
        };
        o << ",\"gn_row\":" << plan.gn_row << ",\"rows\":" << plan.rows << ",\"forward\":";
        ops(plan.forward);
        o << ",\"backward\":";
        ops(plan.backward);
      } else {
        o << ",\"error\":";
        str(e);
      }
      o << "}";
    }
    o << "}";
  }
  o << "],\"fields\":[";
  for (size_t i = 0; i < s.itmon.size(); i++) {
    if (i) o << ",";
    str(s.seg[s.itmon[i]].name);
  }
  o << "],\"solver\":{\"method\":" << s.solver.method << ",\"tolerance\":" << s.solver.tolerance
    << ",\"deltamax\":" << s.solver.deltamax << ",\"m\":" << s.solver.m
    << ",\"iterationlimit\":" << s.solver.iterationlimit << ",\"max_g\":" << (s.solver.max_g ? "true" : "false")
    << "}}";
  return o.str();
}

bool read_guess_text(const std::string &text, const SystemSpec &s, std::vector<double> &x, std::string &err) {
  std::istringstream is(text);
  std::string method, charged;
  int mx, my, mz, fjc, n1, n2;
  if (!(is >> method >> mx >> my >> mz >> fjc >> charged >> n1)) {
    err = "guess file header not understood";
    return false;
  }
  std::vector<std::string> names(n1);
  for (int i = 0; i < n1; i++) is >> names[i];
  if (!(is >> n2)) {
    err = "guess file header not understood";
    return false;
  }
  for (int i = 0; i < n2; i++) {
    std::string t;
    is >> t;
  }
  const LatticeSpec &l = s.lat;
  int m;
  if (mz > 0)
    m = (mx + 2) * (my + 2) * (mz + 2);
  else if (my > 0)
    m = (mx + 2 * fjc) * (my + 2 * fjc);
  else
    m = mx + 2 * fjc;
  if (mx != l.MX || my != l.MY || mz != l.MZ || m != l.M) {
    err = "guess file lattice size differs from the system";
    return false;
  }
  int iv = (n1 + n2) * m + (charged == "true" ? m : 0);
  std::vector<double> vals(iv);
  for (int i = 0; i < iv; i++)
    if (!(is >> vals[i])) {
      err = "guess file has too few values";
      return false;
    }
  x.assign(s.nvar(), 0.0);
  for (int i = 0; i < n1; i++)
    for (size_t j = 0; j < s.itmon.size(); j++)
      if (names[i] == s.seg[s.itmon[j]].name)
        std::copy(vals.begin() + i * m, vals.begin() + (i + 1) * m, x.begin() + j * l.M);
  // the potential block follows the field blocks when both are charged
  if (charged == "true" && s.charged)
    std::copy(vals.begin() + (n1 + n2) * m, vals.begin() + (n1 + n2 + 1) * m, x.begin() + s.itmon.size() * l.M);
  return true;
}

}  // namespace nscf
