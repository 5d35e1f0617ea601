// ORACLE — test infrastructure only.
//
// CPU FP64 restatement of the reference's lattice-SCF hot path, used as the
// checker for the HIP path.  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load it; nothing in namics_amd/ calls it.
//
// Every routine follows the reference's arithmetic order (the element-wise
// primitives of src/tools_host.h:185-522 applied in the same sequence), so
// that at given fields it reproduces the reference to rounding:
//   set_bounds / remove_bounds   src/LGrad1.cpp:773-861, src/LGrad2.cpp:1517-1700,
//                                src/LGrad3.cpp:1523-1677
//   propagate (Markov 1)         src/LG1Planar.cpp:138-159, src/LGrad1.cpp:432-454,
//                                src/LGrad2.cpp:1053-1073, src/LG2Planar.cpp:1039-1108,
//                                src/LGrad3.cpp:673-743
//   propagateF/B (Markov 2, 1-D) src/LG1Planar.cpp:57-136, src/LGrad1.cpp:285-429
//   Side                         src/LG1Planar.cpp:28-55, src/LGrad1.cpp:251-281,
//                                src/LGrad2.cpp:207-278, src/LG2Planar.cpp:27-135,
//                                src/LGrad3.cpp:103-170
//   ComputeGN / WeightedSum /    src/LGrad1.cpp:213-221,943-969,
//   AddPhiS / Initiate           src/LGrad2.cpp:147-157,1814-1900, src/LGrad3.cpp:40-46,1883-1935
//   ComputePhi (linear chains)   src/mol_linear.cpp:131-149, src/molecule.cpp:2300-2615
//   PrepareForCalculations       src/segment.cpp:835-879
//   ComputePhis normalisation    src/system.cpp:2324-2717
//   Classical_residual           src/system.cpp:2087-2189
//   GetFreeEnergy / GetGrandPotential / CreateMu
//                                src/system.cpp:2809-3428
//   iterate_DIIS / DIIS / Ax     src/sfnewton.cpp:793-880,1037-1130
//   svdcmp (Golub-Reinsch SVD as published in Numerical Recipes in C, 2nd ed.,
//   section 2.6, stored transposed as in src/tools.cpp:1076-1349)
//
// save_memory is not restated: the reference's checkpointed propagation
// returns bit-identical densities to full storage, so the oracle always
// stores all end-point distributions.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "../namics_amd/csrc/scf_spec.h"

using namespace nscf;
typedef std::vector<double> vec;

namespace {

// ---------------------------------------------------------------- primitives
inline void Zero(double *p, int M) { std::fill(p, p + M, 0.0); }
inline void Cp(double *p, const double *a, int M) { std::copy(a, a + M, p); }
inline void Add(double *p, const double *a, int M) {
  for (int i = 0; i < M; i++) p[i] += a[i];
}
inline void Times(double *p, const double *a, const double *b, int M) {
  for (int i = 0; i < M; i++) p[i] = a[i] * b[i];
}
inline void AddTimes(double *p, const double *a, const double *b, int M) {
  for (int i = 0; i < M; i++) p[i] += a[i] * b[i];
}
inline void YplusisCtimesX(double *y, const double *x, double c, int M) {
  for (int i = 0; i < M; i++) y[i] += c * x[i];
}
inline void YplusisCtimesAtimesB(double *y, const double *a, const double *b, double c, int M) {
  for (int i = 0; i < M; i++) y[i] += c * a[i] * b[i];
}
inline void Norm(double *p, double c, int M) {
  for (int i = 0; i < M; i++) p[i] *= c;
}
inline void Div(double *p, const double *a, int M) {
  for (int i = 0; i < M; i++)
    if (a[i] != 0)
      p[i] /= a[i];
    else
      p[i] = 0;
}
inline double Sum(const double *x, int M) { return std::accumulate(x, x + M, 0.0); }
inline double Dot(const double *x, const double *y, int M) {
  double r = 0.0;
  for (int i = 0; i < M; i++) r += x[i] * y[i];
  return r;
}

// ----------------------------------------------------------------- lattice
struct Lat {
  const LatticeSpec &l;
  explicit Lat(const LatticeSpec &ls) : l(ls) {}

  void set_bounds(double *X) const {
    const int MX = l.MX, MY = l.MY, MZ = l.MZ, JX = l.JX, JY = l.JY;
    if (l.gradients == 1 && l.fjc > 1) {
      // refined lattice: fjc ghost layers per side (src/LGrad1.cpp:823-835)
      for (int k = 0; k < l.fjc; k++) {
        X[k] = X[l.bx1f[k]];
        X[MX + l.fjc + k] = X[l.bxmf[k]];
      }
    } else if (l.gradients == 1) {
      X[0] = X[l.BX1];
      X[MX + 1] = X[l.BXM];
    } else if (l.gradients == 2 && l.fjc > 1) {
      // src/LGrad2.cpp:1674-1687: y ghosts of the interior x, then x ghosts
      // of every y (corners included)
      const int f = l.fjc;
      for (int x = f; x < MX + f; x++)
        for (int k = 0; k < f; k++) {
          X[x * JX + k] = X[x * JX + l.by1f[k]];
          X[x * JX + MY + f + k] = X[x * JX + l.bymf[k]];
        }
      for (int y = 0; y < MY + 2 * f; y++)
        for (int k = 0; k < f; k++) {
          X[k * JX + y] = X[l.bx1f[k] * JX + y];
          X[(MX + f + k) * JX + y] = X[l.bxmf[k] * JX + y];
        }
    } else if (l.gradients == 2) {
      for (int x = 1; x < MX + 1; x++) {
        X[x * JX + 0] = X[x * JX + l.BY1];
        X[x * JX + MY + 1] = X[x * JX + l.BYM];
      }
      for (int y = 1; y < MY + 1; y++) {
        X[0 + y] = X[l.BX1 * JX + y];
        X[(MX + 1) * JX + y] = X[l.BXM * JX + y];
      }
      X[0] = X[l.BY1];
      X[MY + 1] = X[l.BYM];
      X[(MX + 1) * JX + 0] = X[(MX + 1) * JX + l.BY1];
      X[(MX + 1) * JX + MY + 1] = X[(MX + 1) * JX + l.BYM];
    } else if (l.fjc > 1) {
      // src/LGrad3.cpp:1618-1672: faces of the interior, then the edges of the
      // x-ghost planes and of the y-ghost rows, then the corners
      const int f = l.fjc;
      auto P3 = [&](int x, int y, int z) { return x * JX + y * JY + z; };
      for (int x = f; x < MX + f; x++)
        for (int y = f; y < MY + f; y++)
          for (int k = 0; k < f; k++) {
            X[P3(x, y, k)] = X[P3(x, y, l.bz1f[k])];
            X[P3(x, y, MZ + f + k)] = X[P3(x, y, l.bzmf[k])];
          }
      for (int y = f; y < MY + f; y++)
        for (int z = f; z < MZ + f; z++)
          for (int k = 0; k < f; k++) {
            X[P3(k, y, z)] = X[P3(l.bx1f[k], y, z)];
            X[P3(MX + f + k, y, z)] = X[P3(l.bxmf[k], y, z)];
          }
      for (int z = f; z < MZ + f; z++)
        for (int x = f; x < MX + f; x++)
          for (int k = 0; k < f; k++) {
            X[P3(x, k, z)] = X[P3(x, l.by1f[k], z)];
            X[P3(x, MY + f + k, z)] = X[P3(x, l.bymf[k], z)];
          }
      auto xplane = [&](int x) {
        for (int y = f; y < MY + f; y++)
          for (int k = 0; k < f; k++) {
            X[P3(x, y, k)] = X[P3(x, y, l.bz1f[k])];
            X[P3(x, y, MZ + f + k)] = X[P3(x, y, l.bzmf[k])];
          }
        for (int z = f; z < MZ + f; z++)
          for (int k = 0; k < f; k++) {
            X[P3(x, k, z)] = X[P3(x, l.by1f[k], z)];
            X[P3(x, MY + f + k, z)] = X[P3(x, l.bymf[k], z)];
          }
      };
      for (int x = 0; x < f; x++) xplane(x);
      for (int x = MX + f; x < MX + 2 * f; x++) xplane(x);
      auto yrow = [&](int y) {
        for (int x = f; x < MX + f; x++)
          for (int k = 0; k < f; k++) {
            X[P3(x, y, k)] = X[P3(x, y, l.bz1f[k])];
            X[P3(x, y, MZ + f + k)] = X[P3(x, y, l.bzmf[k])];
          }
      };
      for (int y = 0; y < f; y++) yrow(y);
      for (int y = MY + f; y < MY + 2 * f; y++) yrow(y);
      for (int k = 0; k < f; k++)
        for (int q = 0; q < f; q++)
          for (int m = 0; m < f; m++) {
            X[P3(k, q, m)] = X[P3(l.bx1f[k], l.by1f[q], l.bz1f[m])];
            X[P3(MX + f + k, q, m)] = X[P3(l.bxmf[k], l.by1f[q], l.bz1f[m])];
            X[P3(k, MY + f + q, m)] = X[P3(l.bx1f[k], l.bymf[q], l.bz1f[m])];
            X[P3(k, q, MZ + f + m)] = X[P3(l.bx1f[k], l.by1f[q], l.bzmf[m])];
            X[P3(MX + f + k, MY + f + q, m)] = X[P3(l.bxmf[k], l.bymf[q], l.bz1f[m])];
            X[P3(MX + f + k, q, MZ + f + m)] = X[P3(l.bxmf[k], l.by1f[q], l.bzmf[m])];
            X[P3(k, MY + f + q, MZ + f + m)] = X[P3(l.bx1f[k], l.bymf[q], l.bzmf[m])];
            X[P3(MX + f + k, MY + f + q, MZ + f + m)] = X[P3(l.bxmf[k], l.bymf[q], l.bzmf[m])];
          }
    } else {
      const int JZ = 1;
      const int BX1 = l.BX1, BXM = l.BXM, BY1 = l.BY1, BYM = l.BYM, BZ1 = l.BZ1, BZM = l.BZM;
      for (int x = 1; x < MX + 1; x++)
        for (int y = 1; y < MY + 1; y++) {
          X[x * JX + y * JY + 0] = X[x * JX + y * JY + BZ1];
          X[x * JX + y * JY + MZ + 1] = X[x * JX + y * JY + BZM];
        }
      for (int y = 1; y < MY + 1; y++)
        for (int z = 1; z < MZ + 1; z++) {
          X[0 + y * JY + z * JZ] = X[BX1 * JX + y * JY + z * JZ];
          X[(MX + 1) * JX + y * JY + z * JZ] = X[BXM * JX + y * JY + z * JZ];
        }
      for (int z = 1; z < MZ + 1; z++)
        for (int x = 1; x < MX + 1; x++) {
          X[x * JX + 0 + z * JZ] = X[x * JX + BY1 * JY + z * JZ];
          X[x * JX + (MY + 1) * JY + z * JZ] = X[x * JX + BYM * JY + z * JZ];
        }
      for (int x : {0, MX + 1}) {
        for (int y = 1; y < MY + 1; y++) {
          X[x * JX + y * JY + 0] = X[x * JX + y * JY + BZ1];
          X[x * JX + y * JY + MZ + 1] = X[x * JX + y * JY + BZM];
        }
        for (int z = 1; z < MZ + 1; z++) {
          X[x * JX + 0 + z * JZ] = X[x * JX + BY1 * JY + z * JZ];
          X[x * JX + (MY + 1) * JY + z * JZ] = X[x * JX + BYM * JY + z * JZ];
        }
      }
      for (int y : {0, MY + 1})
        for (int x = 1; x < MX + 1; x++) {
          X[x * JX + y * JY + 0] = X[x * JX + y * JY + BZ1 * JZ];
          X[x * JX + y * JY + (MZ + 1) * JZ] = X[x * JX + y * JY + BZM * JZ];
        }
      X[0] = X[BX1 * JX + BY1 * JY + BZ1 * JZ];
      X[(MX + 1) * JX] = X[BXM * JX + BY1 * JY + BZ1 * JZ];
      X[(MY + 1) * JY] = X[BX1 * JX + BYM * JY + BZ1 * JZ];
      X[(MZ + 1) * JZ] = X[BX1 * JX + BY1 * JY + BZM * JZ];
      X[(MX + 1) * JX + (MY + 1) * JY] = X[BXM * JX + BYM * JY + BZ1 * JZ];
      X[(MX + 1) * JX + (MZ + 1) * JZ] = X[BXM * JX + BY1 * JY + BZM * JZ];
      X[(MY + 1) * JY + (MZ + 1) * JZ] = X[BX1 * JX + BYM * JY + BZM * JZ];
      X[(MX + 1) * JX + (MY + 1) * JY + (MZ + 1) * JZ] = X[BXM * JX + BYM * JY + BZM * JZ];
    }
  }

  // two-array form used for Markov-2 direction components (src/LGrad1.cpp:787-821)
  void set_bounds2(double *X, double *Y) const {
    if (l.fjc > 1) {
      const int f = l.fjc;
      for (int k = 0; k < f; k++) {
        X[k] = Y[l.bx1f[k]];
        X[l.MX + f + k] = Y[l.bxmf[k]];
        Y[k] = X[l.bx1f[k]];
        Y[l.MX + f + k] = X[l.bxmf[k]];
      }
      return;
    }
    X[0] = Y[l.BX1];
    X[l.MX + 1] = Y[l.BXM];
    Y[0] = X[l.BX1];
    Y[l.MX + 1] = X[l.BXM];
  }

  // Markov-2 step on a planar 1-gradient hexagonal lattice (any fjc), FJC
  // direction components: component p moves fjc - p sites per bond; the
  // weight from component q is P[|p-q|] plus, for q not along x,
  // P[FJC-1-|FJC-1-p-q|] (LG1Planar::propagateF/B src/LG1Planar.cpp:82-94,
  // 122-134).
  void propagate_hex(double *gs, double *gs_1, const double *G1, const double *P, bool forward) const {
    const int M = l.M, f = l.fjc, F = l.FJC;
    Zero(gs, M * F);
    for (int k = 0; k < (F - 1) / 2; k++) set_bounds2(gs_1 + k * M, gs_1 + (F - k - 1) * M);
    set_bounds(gs_1 + (F - 1) / 2 * M);
    auto cw = [&](int p, int q) {
      double c = P[std::abs(-p + q)];
      if (q > 0 && q < F - 1) c += P[F - 1 - std::abs(F - 1 - p - q)];
      return c;
    };
    if (l.geometry != GEOM_PLANAR) {
      // curved hexagonal lattice, fjc = 1 (LGrad1::propagateF / propagateB
      // src/LGrad1.cpp:298-321,374-398): a step off the axis mixes the
      // component with its reflection, with the Markov-2 curvature weights
      vec Hv(M, 0.0);
      double *H = Hv.data();
      const double *l1 = l.l1.data(), *l11 = l.l11.data(), *l_1 = l.l_1.data(), *l_11 = l.l_11.data();
      auto fillH = [&](int q, int a, int b) {
        if (a > 0) {
          Times(H + b, l_1 + a, gs_1 + q * M + b, M - a - b);
          AddTimes(H + b, l_11 + a, gs_1 + (F - 1 - q) * M + a, M - a - b);
        }
        if (b > 0) {
          Times(H + b, l1 + a, gs_1 + q * M + b, M - a - b);
          AddTimes(H + b, l11 + a, gs_1 + (F - 1 - q) * M + a, M - a - b);
        }
      };
      auto addp = [&](int p, int q, int a, int b) {
        const double c = cw(p, q);
        if (a + b > 0) {
          if (c != 0) YplusisCtimesX(gs + p * M + a, H + b, c, M - a - b);
        } else {
          if (c != 0) YplusisCtimesX(gs + p * M, gs_1 + q * M, c, M);
        }
      };
      if (forward) {
        for (int p = 0; p < F; p++) {
          int a = p - f, b;
          if (a < 0) {
            b = 0;
            a = -a;
          } else {
            b = a;
            a = 0;
          }
          for (int q = 0; q < F; q++) {
            fillH(q, a, b);
            addp(p, q, a, b);
          }
        }
      } else {
        for (int q = F - 1; q > -1; q--) {
          int a = q - f, b;
          if (a > 0) {
            b = 0;
          } else {
            b = -a;
            a = 0;
          }
          fillH(q, a, b);
          for (int p = F - 1; p > -1; p--) addp(p, q, a, b);
        }
      }
      for (int k = 0; k < F; k++) Times(gs + k * M, gs + k * M, G1, M);
      return;
    }
    if (forward) {
      for (int p = 0; p < F; p++) {
        int a = p - f, b;
        if (a < 0) {
          b = 0;
          a = -a;
        } else {
          b = a;
          a = 0;
        }
        for (int q = 0; q < F; q++) {
          const double c = cw(p, q);
          if (c != 0) YplusisCtimesX(gs + p * M + a, gs_1 + q * M + b, c, M - a - b);
        }
      }
    } else {
      for (int q = F - 1; q > -1; q--) {
        int a = q - f, b;
        if (a > 0) {
          b = 0;
        } else {
          b = -a;
          a = 0;
        }
        for (int p = F - 1; p > -1; p--) {
          const double c = cw(p, q);
          if (c != 0) YplusisCtimesX(gs + p * M + a, gs_1 + q * M + b, c, M - a - b);
        }
      }
    }
    for (int k = 0; k < F; k++) Times(gs + k * M, gs + k * M, G1, M);
  }

  void remove_bounds(double *X) const {
    const int MX = l.MX, MY = l.MY, MZ = l.MZ, JX = l.JX, JY = l.JY;
    if (l.gradients == 1 && l.fjc > 1) {
      for (int k = 0; k < l.fjc; k++) {
        X[k] = 0;
        X[MX + l.fjc + k] = 0;
      }
    } else if (l.gradients == 1) {
      X[0] = 0;
      X[MX + 1] = 0;
    } else if (l.gradients == 2 && l.fjc > 1) {
      const int f = l.fjc;
      for (int x = f; x < MX + f; x++)
        for (int k = 0; k < f; k++) {
          X[x * JX + k] = 0;
          X[x * JX + MY + f + k] = 0;
        }
      for (int y = 0; y < MY + 2 * f; y++)
        for (int k = 0; k < f; k++) {
          X[k * JX + y] = 0;
          X[(MX + f + k) * JX + y] = 0;
        }
    } else if (l.gradients == 2) {
      for (int x = 1; x < MX + 1; x++) {
        X[x * JX + 0] = 0;
        X[x * JX + MY + 1] = 0;
      }
      for (int y = 1; y < MY + 1; y++) {
        X[0 + y] = 0;
        X[(MX + 1) * JX + y] = 0;
      }
      X[0] = 0;
      X[MY + 1] = 0;
      X[(MX + 1) * JX + 0] = 0;
      X[(MX + 1) * JX + MY + 1] = 0;
    } else if (l.fjc > 1) {
      const int f = l.fjc;
      for (int x = 0; x < MX + 2 * f; x++)
        for (int y = 0; y < MY + 2 * f; y++)
          for (int k = 0; k < f; k++) {
            X[x * JX + y * JY + k] = 0;
            X[x * JX + y * JY + MZ + f + k] = 0;
          }
      for (int y = 0; y < MY + 2 * f; y++)
        for (int z = 0; z < MZ + 2 * f; z++)
          for (int k = 0; k < f; k++) {
            X[k * JX + y * JY + z] = 0;
            X[(MX + f + k) * JX + y * JY + z] = 0;
          }
      for (int z = 0; z < MZ + 2 * f; z++)
        for (int x = 0; x < MX + 2 * f; x++)
          for (int k = 0; k < f; k++) {
            X[x * JX + k * JY + z] = 0;
            X[x * JX + (MY + f + k) * JY + z] = 0;
          }
    } else {
      // RemoveBoundaries (src/tools_host.h:305-310): zero every ghost face
      for (int y = 0; y < MY + 2; y++)
        for (int z = 0; z < MZ + 2; z++) {
          X[y * JY + z] = 0;
          X[(MX + 1) * JX + y * JY + z] = 0;
        }
      for (int x = 0; x < MX + 2; x++)
        for (int z = 0; z < MZ + 2; z++) {
          X[x * JX + z] = 0;
          X[x * JX + (MY + 1) * JY + z] = 0;
        }
      for (int x = 0; x < MX + 2; x++)
        for (int y = 0; y < MY + 2; y++) {
          X[x * JX + y * JY] = 0;
          X[x * JX + y * JY + MZ + 1] = 0;
        }
    }
  }

  double weighted_sum(double *X) const {
    remove_bounds(X);
    const int M = l.M;
    if (l.gradients == 3) return Sum(X, M);
    if (l.geometry == GEOM_PLANAR && l.gradients == 2) return l.fjc > 1 ? Sum(X, M) / (l.fjc * l.fjc) : Sum(X, M);
    if (l.geometry == GEOM_PLANAR) return l.fjc > 1 ? Sum(X, M) / l.fjc : Sum(X, M);
    return Dot(X, l.L.data(), M);
  }

  // Refined 1-gradient stencil (fjc > 1): FJC taps with the LAMBDA rows,
  // in the order of src/LG1Planar.cpp:151-156 / src/LGrad1.cpp:447-452
  // (also Side, src/LG1Planar.cpp:47-52, src/LGrad1.cpp:272-277).
  void refined_taps(double *out, const double *X) const {
    const int M = l.M, f = l.fjc, F = l.FJC;
    const double *LA = l.LAMBDA.data();
    for (int j = 0; j < F / 2; j++) {
      const int kk = (F - 1) / 2 - j;
      AddTimes(out + kk, X, LA + j * M + kk, M - kk);
      AddTimes(out, X + kk, LA + (F - j - 1) * M, M - kk);
    }
    AddTimes(out, X, LA + f * M, M);
  }

  // 3-gradient stencils over the 26 neighbours, the passes of
  // LGrad3::propagate / Side without stencil_full (src/LGrad3.cpp:686-729,
  // 111-152; simple cubic 4, 4, 1/152; hexagonal 2, 2, 1/56) and the
  // hexagonal stencil_full form (src/LGrad3.cpp:746-782, 166-205), with the
  // reference's pass lengths.
  void wide3(double *gs, double *gs_1) const {
    const int M = l.M, JX = l.JX, JY = l.JY;
    const bool sc = l.lattice_type == LAT_SIMPLE_CUBIC;
    if (!l.stencil_full) {
      const double n1 = sc ? 4.0 : 2.0;
      Add(gs + JX, gs_1, M - JX);
      Add(gs, gs_1 + JX, M - JX);
      Add(gs + JY, gs_1, M - JY);
      Add(gs, gs_1 + JY, M - JY);
      Add(gs + 1, gs_1, M - 1);
      Add(gs, gs_1 + 1, M - 1);
      Norm(gs, n1, M);
      Add(gs + JX + JY, gs_1, M - JX - JY);
      Add(gs, gs_1 + JX + JY, M - JX - JY);
      Add(gs + JY, gs_1 + JX, M - JY - JX);
      Add(gs + JX, gs_1 + JY, M - JY - JX);
      Add(gs + JX + 1, gs_1, M - JX - 1);
      Add(gs, gs_1 + JX + 1, M - JX - 1);
      Add(gs + JX, gs_1 + 1, M - JX);
      Add(gs + 1, gs_1 + JX, M - JX);
      Add(gs + JY + 1, gs_1, M - JY - 1);
      Add(gs, gs_1 + JY + 1, M - JX - 1);
      Add(gs + JY, gs_1 + 1, M - JY);
      Add(gs + 1, gs_1 + JY, M - JY);
      Norm(gs, n1, M);
    } else {
      const double Two = 2.0;
      YplusisCtimesX(gs, gs_1, Two, M);
      Add(gs + JX, gs_1, M - JX);
      Add(gs, gs_1 + JX, M - JX);
      Add(gs + JY, gs_1, M - JY);
      Add(gs, gs_1 + JY, M - JY);
      Add(gs + 1, gs_1, M - 1);
      Add(gs, gs_1 + 1, M - 1);
      Norm(gs, Two, M);
      Add(gs + JX + JY, gs_1, M - JX - JY);
      Add(gs, gs_1 + JX + JY, M - JX - JY);
      Add(gs + JX, gs_1 + JY, M - JX - JY);
      Add(gs + JY, gs_1 + JX, M - JX - JY);
      Add(gs + JX + 1, gs_1, M - JX - 1);
      Add(gs + 1, gs_1 + JX, M - JX - 1);
      Add(gs + JY + 1, gs_1, M - JY - 1);
      Add(gs + 1, gs_1 + JY, M - JY - 1);
      Add(gs + JX, gs_1 + 1, M - JX - 1);
      Add(gs, gs_1 + JX + 1, M - JX - 1);
      Add(gs + JY, gs_1 + 1, M - JY - 1);
      Add(gs, gs_1 + JY + 1, M - JY - 1);
      Norm(gs, Two, M);
    }
    Add(gs + JX + JY + 1, gs_1, M - JX - JY - 1);
    Add(gs, gs_1 + JX + JY + 1, M - JX - JY - 1);
    Add(gs + JX + JY, gs_1 + 1, M - JX - JY - 1);
    Add(gs + 1, gs_1 + JX + JY, M - JX - JY - 1);
    Add(gs + JX + 1, gs_1 + JY, M - JX - JY - 1);
    Add(gs + JY, gs_1 + JX + 1, M - JX - JY - 1);
    Add(gs + JY + 1, gs_1 + JX, M - JX - JY - 1);
    Add(gs + JX, gs_1 + JY + 1, M - JX - JY - 1);
    Norm(gs, l.stencil_full ? 1.0 / 40.0 : (sc ? 1.0 / 152.0 : 1.0 / 56.0), M);
  }

  // Refined 2-gradient stencils (fjc > 1), block sums over the
  // (2 fjc + 1)^2 offsets.  Cylindrical (LGrad2::propagate / Side
  // src/LGrad2.cpp:1091-1106, 264-278): offsets with |dy| < fjc, weighted
  // by the LAMBDA row of dx, times 2, plus those with |dy| = fjc, times
  // 1/(2(fjc+2)).  Planar (LG2Planar src/LG2Planar.cpp:1110-1124, 117-130):
  // plain sums of the offsets with no, one, two coordinates at +-fjc, with
  // factors 2, 2 and 1/(4(FJC-2)FJC+1).
  void refined2(double *gs, double *gs_1) const {
    const int M = l.M, JX = l.JX, JY = l.JY, f = l.fjc, F = l.FJC;
    const double *LA = l.LAMBDA.data();
    const bool planar = l.geometry == GEOM_PLANAR;
    const int nblocks = planar ? 3 : 2;
    for (int block = 0; block < nblocks; block++) {
      for (int x = -f; x < f + 1; x++)
        for (int y = -f; y < f + 1; y++) {
          int bk = 0, a = 0, b = 0;
          if (planar && (x == -f || x == f)) bk++;
          if (y == -f || y == f) bk++;
          if (bk != block) continue;
          if (x < 0) a = -x * JX; else b = x * JX;
          if (y < 0) a -= y * JY; else b += y * JY;
          if (planar)
            Add(gs + a, gs_1 + b, M - a - b);
          else
            AddTimes(gs + a, gs_1 + b, LA + (f + x) * M + a, M - a - b);
        }
      if (planar)
        Norm(gs, block != 2 ? 2.0 : 1.0 / (4.0 * (F - 2) * F + 1), M);
      else
        Norm(gs, block != 1 ? 2.0 : 1.0 / (2 * (f + 2)), M);
    }
  }

  // Refined 3-gradient hexagonal stencil (fjc = 2; LGrad3::propagate / Side
  // src/LGrad3.cpp:808-830, 206-226): plain sums of the (2 fjc + 1)^3
  // offsets with none, one, two, three coordinates at +-fjc, with factors 2,
  // 2, 2 and 1/8/((FJC-2)^3 + 3(FJC-2)^2 + 3(FJC-2) + 1).
  void refined3(double *gs, double *gs_1) const {
    const int M = l.M, JX = l.JX, JY = l.JY, f = l.fjc, F = l.FJC;
    for (int block = 0; block < 4; block++) {
      for (int x = -f; x < f + 1; x++)
        for (int y = -f; y < f + 1; y++)
          for (int z = -f; z < f + 1; z++) {
            int bk = 0, a = 0, b = 0;
            if (x == -f || x == f) bk++;
            if (y == -f || y == f) bk++;
            if (z == -f || z == f) bk++;
            if (bk != block) continue;
            if (x < 0) a = -x * JX; else b = x * JX;
            if (y < 0) a -= y * JY; else b += y * JY;
            if (z < 0) a -= z; else b += z;
            Add(gs + a, gs_1 + b, M - a - b);
          }
      Norm(gs, block != 3 ? 2.0 : 1.0 / 8.0 / ((F - 2) * (F - 2) * (F - 2) + 3 * (F - 2) * (F - 2) + 3 * (F - 2) + 1),
           M);
    }
  }

  // Markov-1 step: gs <- G1 * W(gs_1); gs_1 gets its ghosts set in place.
  void propagate(double *gs, double *gs_1, const double *G1) const {
    const int M = l.M, JX = l.JX;
    Zero(gs, M);
    set_bounds(gs_1);
    if (l.gradients == 1 && l.fjc > 1) {
      refined_taps(gs, gs_1);
      Times(gs, gs, G1, M);
    } else if (l.gradients == 2 && l.fjc > 1) {
      refined2(gs, gs_1);
      Times(gs, gs, G1, M);
    } else if (l.gradients == 1) {
      if (l.geometry == GEOM_PLANAR) {
        const double lam = l.lambda;
        YplusisCtimesX(gs + 1, gs_1, lam, M - 1);
        YplusisCtimesX(gs, gs_1, 1.0 - 2.0 * lam, M);
        YplusisCtimesX(gs, gs_1 + 1, lam, M - 1);
      } else {
        AddTimes(gs, gs_1, l.lambda0.data(), M);
        AddTimes(gs + 1, gs_1, l.lambda_1.data() + 1, M - 1);
        AddTimes(gs, gs_1 + 1, l.lambda1.data(), M - 1);
      }
      Times(gs, gs, G1, M);
    } else if (l.gradients == 2) {
      if (l.geometry == GEOM_PLANAR) {
        if (l.stencil_full) {
          if (l.lattice_type == LAT_SIMPLE_CUBIC) {
            const double C1 = 16.0 / 36.0, C2 = 4.0 / 36.0, C3 = 1.0 / 36.0;
            YplusisCtimesX(gs, gs_1, C1, M);
            YplusisCtimesX(gs + 1, gs_1, C2, M - 1);
            YplusisCtimesX(gs, gs_1 + 1, C2, M - 1);
            YplusisCtimesX(gs + JX, gs_1, C2, M - JX);
            YplusisCtimesX(gs, gs_1 + JX, C2, M - JX);
            YplusisCtimesX(gs + JX + 1, gs_1, C3, M - JX - 1);
            YplusisCtimesX(gs + JX, gs_1 + 1, C3, M - JX);
            YplusisCtimesX(gs + 1, gs_1 + JX, C3, M - JX);
            YplusisCtimesX(gs, gs_1 + JX + 1, C3, M - JX - 1);
          } else {
            const double Two = 2.0, C = 1.0 / 16.0;
            const int JY = l.JY;
            YplusisCtimesX(gs, gs_1, Two, M);
            Add(gs + JX, gs_1, M - JX);
            Add(gs, gs_1 + JX, M - JX);
            Add(gs + JY, gs_1, M - JY);
            Add(gs, gs_1 + JY, M - JY);
            Norm(gs, Two, M);
            Add(gs + JX + JY, gs_1, M - JX - JY);
            Add(gs, gs_1 + JX + JY, M - JX - JY);
            Add(gs + JX, gs_1 + JY, M - JX - JY);
            Add(gs + JY, gs_1 + JX, M - JX - JY);
            Norm(gs, C, M);
          }
        } else {
          if (l.lattice_type != LAT_SIMPLE_CUBIC) throw std::string("unsupported 2-D stencil");
          const int JY = l.JY;
          Add(gs + JX, gs_1, M - JX);
          Add(gs, gs_1 + JX, M - JX);
          Add(gs + JY, gs_1, M - JY);
          Add(gs, gs_1 + JY, M - JY);
          Norm(gs, 1.0 / 2.0, M);
          Add(gs, gs_1, M);
          Norm(gs, 1.0 / 3.0, M);
        }
      } else {
        const bool sc = l.lattice_type == LAT_SIMPLE_CUBIC;
        const double C1 = sc ? 4.0 / 6.0 : 0.5, C2 = sc ? 1.0 / 6.0 : 0.25, C3 = sc ? 4.0 : 2.0;
        const double *lm = l.lambda_1.data(), *lp = l.lambda1.data();
        YplusisCtimesX(gs, gs_1, C1, M);
        AddTimes(gs + JX, gs_1, lm + JX, M - JX);
        AddTimes(gs, gs_1 + JX, lp, M - JX);
        YplusisCtimesX(gs + 1, gs_1, C2, M - 1);
        YplusisCtimesX(gs, gs_1 + 1, C2, M - 1);
        Norm(gs, C3, M);
        AddTimes(gs + JX + 1, gs_1, lm + JX + 1, M - JX - 1);
        AddTimes(gs + JX, gs_1 + 1, lm + JX, M - JX - 1);
        AddTimes(gs + 1, gs_1 + JX, lp + 1, M - JX - 1);
        AddTimes(gs, gs_1 + JX + 1, lp, M - JX - 1);
        Norm(gs, C2, M);
      }
      Times(gs, gs, G1, M);
    } else {
      const int JY = l.JY;
      if (l.fjc > 1) {
        refined3(gs, gs_1);
        Times(gs, gs, G1, M);
        return;
      }
      if (!(l.stencil_full && l.lattice_type == LAT_SIMPLE_CUBIC)) {
        wide3(gs, gs_1);
        Times(gs, gs, G1, M);
        return;
      }
      Add(gs + JX, gs_1, M - JX);
      Add(gs, gs_1 + JX, M - JX);
      Add(gs + JY, gs_1, M - JY);
      Add(gs, gs_1 + JY, M - JY);
      Add(gs + 1, gs_1, M - 1);
      Add(gs, gs_1 + 1, M - 1);
      Norm(gs, 1.0 / 6.0, M);
      Times(gs, gs, G1, M);
    }
  }

  // Markov-2 forward/backward steps on a 1-gradient simple-cubic lattice.
  // gs, gs_1 hold the three direction components back to back (stride M).
  // 2-gradient Markov-2 bounds (LGrad2::set_bounds_x/y src/LGrad2.cpp:1555-1650,
  // fjc = 1, no shift): a pair of opposite directions swaps at a reflecting
  // boundary, every other component is bounded by itself; periodic x or y
  // bounds each component by itself.  Ghost corners are not touched.
  void sbx(double *X) const {
    const int MX = l.MX, MY = l.MY, JX = l.JX;
    for (int y = 1; y < MY + 1; y++) {
      X[0 + y] = X[l.BX1 * JX + y];
      X[(MX + 1) * JX + y] = X[l.BXM * JX + y];
    }
  }
  void sbx(double *X, double *Y) const {
    const int MX = l.MX, MY = l.MY, JX = l.JX;
    if (l.BX1 > l.BXM) {
      sbx(X);
      sbx(Y);
      return;
    }
    for (int y = 1; y < MY + 1; y++) {
      X[0 + y] = Y[l.BX1 * JX + y];
      X[(MX + 1) * JX + y] = Y[l.BXM * JX + y];
      Y[0 + y] = X[l.BX1 * JX + y];
      Y[(MX + 1) * JX + y] = X[l.BXM * JX + y];
    }
  }
  void sby(double *X) const {
    const int MX = l.MX, MY = l.MY, JX = l.JX;
    for (int x = 1; x < MX + 1; x++) {
      X[x * JX + 0] = X[x * JX + l.BY1];
      X[x * JX + MY + 1] = X[x * JX + l.BYM];
    }
  }
  void sby(double *X, double *Y) const {
    const int MX = l.MX, MY = l.MY, JX = l.JX;
    if (l.BY1 > l.BYM) {
      sby(X);
      sby(Y);
      return;
    }
    for (int x = 1; x < MX + 1; x++) {
      X[x * JX + 0] = Y[x * JX + l.BY1];
      X[x * JX + MY + 1] = Y[x * JX + l.BYM];
      Y[x * JX + 0] = X[x * JX + l.BY1];
      Y[x * JX + MY + 1] = X[x * JX + l.BYM];
    }
  }

  // 2-gradient Markov-2 step, simple cubic, 5 direction components (+x, +y,
  // out of plane, -y, -x): LGrad2::propagateF/B (src/LGrad2.cpp:645-681,
  // 1003-1039, with the reflection weights l_1, l_11, l1, l11 of the curved
  // lattice) and LG2Planar::propagateF/B (src/LG2Planar.cpp:445-481,
  // 930-966).
  void propagate2d(double *gs, double *gs_1, const double *G1, const double *P, bool forward) const {
    const int M = l.M, JX = l.JX, JY = l.JY;
    double *z0 = gs_1, *z1 = gs_1 + M, *z2 = gs_1 + 2 * M, *z3 = gs_1 + 3 * M, *z4 = gs_1 + 4 * M;
    sbx(z0, z4);
    sbx(z1);
    sbx(z2);
    sbx(z3);
    sby(z1, z3);
    sby(z0);
    sby(z2);
    sby(z4);
    double *x0 = gs, *x1 = gs + M, *x2 = gs + 2 * M, *x3 = gs + 3 * M, *x4 = gs + 4 * M;
    Zero(gs, 5 * M);
    const bool cyl = l.geometry != GEOM_PLANAR;
    vec Hv(M, 0.0);
    double *H = Hv.data();
    const double *l1 = cyl ? l.l1.data() : nullptr, *l11 = cyl ? l.l11.data() : nullptr;
    const double *l_1 = cyl ? l.l_1.data() : nullptr, *l_11 = cyl ? l.l_11.data() : nullptr;
    // LReflect(H, P, Q): H = l_1(+JX) P + l_11(+JX) Q(+JX); UReflect(H, P, Q):
    // H(+JX) = l1 P(+JX) + l11 Q
    auto lref = [&](const double *Pp, const double *Qq) {
      Times(H, l_1 + JX, Pp, M - JX);
      AddTimes(H, l_11 + JX, Qq + JX, M - JX);
    };
    auto uref = [&](const double *Pp, const double *Qq) {
      Times(H + JX, l1, Pp + JX, M - JX);
      AddTimes(H + JX, l11, Qq, M - JX);
    };
    if (forward) {
      if (cyl) {
        lref(z0, z4);
        YplusisCtimesX(x0 + JX, H, P[0], M - JX);
      } else {
        YplusisCtimesX(x0 + JX, z0, P[0], M - JX);
      }
      YplusisCtimesX(x0 + JX, z1, P[1], M - JX);
      YplusisCtimesX(x0 + JX, z2, 2 * P[1], M - JX);
      YplusisCtimesX(x0 + JX, z3, P[1], M - JX);
      YplusisCtimesX(x1 + JY, z0, P[1], M - JY);
      YplusisCtimesX(x1 + JY, z1, P[0], M - JY);
      YplusisCtimesX(x1 + JY, z2, 2 * P[1], M - JY);
      YplusisCtimesX(x1 + JY, z4, P[1], M - JY);
      YplusisCtimesX(x2, z0, P[1], M);
      YplusisCtimesX(x2, z1, P[1], M);
      YplusisCtimesX(x2, z2, P[0], M);
      YplusisCtimesX(x2, z3, P[1], M);
      YplusisCtimesX(x2, z4, P[1], M);
      YplusisCtimesX(x3, z0 + JY, P[1], M - JY);
      YplusisCtimesX(x3, z2 + JY, 2 * P[1], M - JY);
      YplusisCtimesX(x3, z3 + JY, P[0], M - JY);
      YplusisCtimesX(x3, z4 + JY, P[1], M - JY);
      YplusisCtimesX(x4, z1 + JX, P[1], M - JX);
      YplusisCtimesX(x4, z2 + JX, 2 * P[1], M - JX);
      YplusisCtimesX(x4, z3 + JX, P[1], M - JX);
      if (cyl) {
        uref(z4, z0);
        YplusisCtimesX(x4, H + JX, P[0], M - JX);
      } else {
        YplusisCtimesX(x4, z4 + JX, P[0], M - JX);
      }
    } else {
      const double *h4 = z4;
      if (cyl) {
        lref(z4, z0);
        h4 = H;
      }
      YplusisCtimesX(x1 + JX, h4, P[1], M - JX);
      YplusisCtimesX(x2 + JX, h4, P[1], M - JX);
      YplusisCtimesX(x3 + JX, h4, P[1], M - JX);
      YplusisCtimesX(x4 + JX, h4, P[0], M - JX);
      YplusisCtimesX(x0 + JY, z3, P[1], M - JY);
      YplusisCtimesX(x2 + JY, z3, P[1], M - JY);
      YplusisCtimesX(x3 + JY, z3, P[0], M - JY);
      YplusisCtimesX(x4 + JY, z3, P[1], M - JY);
      YplusisCtimesX(x0, z2, 2 * P[1], M);
      YplusisCtimesX(x1, z2, 2 * P[1], M);
      YplusisCtimesX(x2, z2, P[0], M);
      YplusisCtimesX(x3, z2, 2 * P[1], M);
      YplusisCtimesX(x4, z2, 2 * P[1], M);
      YplusisCtimesX(x0, z1 + JY, P[1], M - JY);
      YplusisCtimesX(x1, z1 + JY, P[0], M - JY);
      YplusisCtimesX(x2, z1 + JY, P[1], M - JY);
      YplusisCtimesX(x4, z1 + JY, P[1], M - JY);
      const double *h0 = z0 + JX;
      if (cyl) {
        uref(z0, z4);
        h0 = H + JX;
      }
      YplusisCtimesX(x0, h0, P[0], M - JX);
      YplusisCtimesX(x1, h0, P[1], M - JX);
      YplusisCtimesX(x2, h0, P[1], M - JX);
      YplusisCtimesX(x3, h0, P[1], M - JX);
    }
    for (int k = 0; k < 5; k++) Times(gs + k * M, gs + k * M, G1, M);
  }

  // 3-gradient Markov-2 face bounds (LGrad3::set_bounds_x/y/z
  // src/LGrad3.cpp:1376-1522, fjc = 1, no shift): axis d's pair of opposite
  // directions swaps at a reflecting boundary of that axis; periodic bounds
  // and the other components take their own values.  Only face ghosts.
  void sb3(double *X, double *Y, int axis) const {
    const int MX = l.MX, MY = l.MY, MZ = l.MZ, JX = l.JX, JY = l.JY;
    const int B1 = axis == 0 ? l.BX1 : (axis == 1 ? l.BY1 : l.BZ1);
    const int BM = axis == 0 ? l.BXM : (axis == 1 ? l.BYM : l.BZM);
    const int n = axis == 0 ? MX : (axis == 1 ? MY : MZ);
    const int J = axis == 0 ? JX : (axis == 1 ? JY : 1);
    const bool pair = Y != nullptr && B1 <= BM;
    // the two other axes span their interior
    const int na = axis == 0 ? MY : MX, nb = axis == 2 ? MY : MZ;
    const int Ja = axis == 0 ? JY : JX, Jb = axis == 2 ? JY : 1;
    for (int a = 1; a < na + 1; a++)
      for (int b = 1; b < nb + 1; b++) {
        const int base = a * Ja + b * Jb;
        if (pair) {
          X[base] = Y[base + B1 * J];
          X[base + (n + 1) * J] = Y[base + BM * J];
          Y[base] = X[base + B1 * J];
          Y[base + (n + 1) * J] = X[base + BM * J];
        } else {
          X[base] = X[base + B1 * J];
          X[base + (n + 1) * J] = X[base + BM * J];
          if (Y) {
            Y[base] = Y[base + B1 * J];
            Y[base + (n + 1) * J] = Y[base + BM * J];
          }
        }
      }
  }

  // 3-gradient Markov-2 step, simple cubic, 6 direction components (+x, +y,
  // +z, -z, -y, -x): LGrad3::propagateF/B (src/LGrad3.cpp:399-446, 605-663);
  // a walker never reverses, straight on weighs P0, a turn P1.
  void propagate3d(double *gs, double *gs_1, const double *G1, const double *P, bool forward) const {
    const int M = l.M, JX = l.JX, JY = l.JY, JZ = 1;
    double *z[6];
    double *o[6];
    for (int k = 0; k < 6; k++) {
      z[k] = gs_1 + k * M;
      o[k] = gs + k * M;
    }
    Zero(gs, 6 * M);
    if (forward) {
      sb3(z[0], z[5], 0);
      for (int k : {1, 2, 3, 4}) sb3(z[k], nullptr, 0);
      sb3(z[1], z[4], 1);
      for (int k : {0, 2, 3, 5}) sb3(z[k], nullptr, 1);
      sb3(z[2], z[3], 2);
      for (int k : {0, 1, 4, 5}) sb3(z[k], nullptr, 2);
      // component p arrives from the site behind it and takes every incoming
      // direction except its reverse 5 - p
      const int sh[6] = {JX, JY, JZ, -JZ, -JY, -JX};
      for (int p = 0; p < 6; p++) {
        const int a = sh[p] > 0 ? sh[p] : 0, b = sh[p] < 0 ? -sh[p] : 0;
        for (int q = 0; q < 6; q++) {
          if (q == 5 - p) continue;
          YplusisCtimesX(o[p] + a, z[q] + b, q == p ? P[0] : P[1], M - a - b);
        }
      }
    } else {
      sb3(z[0], z[5], 0);
      sb3(z[1], z[4], 1);
      sb3(z[2], z[3], 2);
      // incoming direction q (reverse order 5..0) feeds every component p
      // except its reverse
      const int sh[6] = {JX, JY, JZ, -JZ, -JY, -JX};
      for (int q = 5; q >= 0; q--) {
        const int a = sh[5 - q] > 0 ? sh[5 - q] : 0, b = sh[5 - q] < 0 ? -sh[5 - q] : 0;
        for (int p = 0; p < 6; p++) {
          if (p == 5 - q) continue;
          YplusisCtimesX(o[p] + a, z[q] + b, q == p ? P[0] : P[1], M - a - b);
        }
      }
    }
    for (int k = 0; k < 6; k++) Times(gs + k * M, gs + k * M, G1, M);
  }

  void propagateF(double *gs, double *gs_1, const double *G1, const double *P) const {
    if (l.gradients == 3) return propagate3d(gs, gs_1, G1, P, true);
    if (l.gradients == 2) return propagate2d(gs, gs_1, G1, P, true);
    if (l.lattice_type == LAT_HEXAGONAL) return propagate_hex(gs, gs_1, G1, P, true);
    const int M = l.M;
    Zero(gs, 3 * M);
    set_bounds2(gs_1, gs_1 + 2 * M);
    set_bounds(gs_1 + M);
    double *gz0 = gs_1, *gz1 = gs_1 + M, *gz2 = gs_1 + 2 * M;
    double *gx0 = gs, *gx1 = gs + M, *gx2 = gs + 2 * M;
    if (l.geometry == GEOM_PLANAR) {
      YplusisCtimesX(gx0 + 1, gz0, P[0], M - 1);
      YplusisCtimesX(gx0 + 1, gz1, 4 * P[1], M - 1);
      YplusisCtimesX(gx1, gz0, P[1], M);
      YplusisCtimesX(gx1, gz1, 2 * P[1] + P[0], M);
      YplusisCtimesX(gx1, gz2, P[1], M);
      YplusisCtimesX(gx2, gz1 + 1, 4 * P[1], M - 1);
      YplusisCtimesX(gx2, gz2 + 1, P[0], M - 1);
    } else {
      vec Hv(M, 0.0);
      double *H = Hv.data();
      const double *l1 = l.l1.data(), *l11 = l.l11.data(), *l_1 = l.l_1.data(), *l_11 = l.l_11.data();
      Times(H, l_1 + 1, gz0, M - 1);
      AddTimes(H, l_11 + 1, gz2 + 1, M - 1);
      YplusisCtimesX(gx0 + 1, H, P[0], M - 1);
      Times(H, l_1 + 1, gz1, M - 1);
      AddTimes(H, l_11 + 1, gz1 + 1, M - 1);
      YplusisCtimesX(gx0 + 1, H, 4 * P[1], M - 1);
      YplusisCtimesX(gx1, gz0, P[1], M);
      YplusisCtimesX(gx1, gz1, 2 * P[1] + P[0], M);
      YplusisCtimesX(gx1, gz2, P[1], M);
      Times(H + 1, l1, gz1 + 1, M - 1);
      AddTimes(H + 1, l11, gz1, M - 1);
      YplusisCtimesX(gx2, H + 1, 4 * P[1], M - 1);
      Times(H + 1, l1, gz2 + 1, M - 1);
      AddTimes(H + 1, l11, gz0, M - 1);
      YplusisCtimesX(gx2, H + 1, P[0], M - 1);
    }
    for (int k = 0; k < 3; k++) Times(gs + k * M, gs + k * M, G1, M);
  }

  void propagateB(double *gs, double *gs_1, const double *G1, const double *P) const {
    if (l.gradients == 3) return propagate3d(gs, gs_1, G1, P, false);
    if (l.gradients == 2) return propagate2d(gs, gs_1, G1, P, false);
    if (l.lattice_type == LAT_HEXAGONAL) return propagate_hex(gs, gs_1, G1, P, false);
    const int M = l.M;
    Zero(gs, 3 * M);
    set_bounds2(gs_1, gs_1 + 2 * M);
    set_bounds(gs_1 + M);
    double *gz0 = gs_1, *gz1 = gs_1 + M, *gz2 = gs_1 + 2 * M;
    double *gx0 = gs, *gx1 = gs + M, *gx2 = gs + 2 * M;
    if (l.geometry == GEOM_PLANAR) {
      YplusisCtimesX(gx0, gz0 + 1, P[0], M - 1);
      YplusisCtimesX(gx0, gz1, 4 * P[1], M);
      YplusisCtimesX(gx1, gz0 + 1, P[1], M - 1);
      YplusisCtimesX(gx1, gz1, 2 * P[1] + P[0], M);
      YplusisCtimesX(gx1 + 1, gz2, P[1], M - 1);
      YplusisCtimesX(gx2, gz1, 4 * P[1], M);
      YplusisCtimesX(gx2 + 1, gz2, P[0], M - 1);
    } else {
      vec Hv(M, 0.0);
      double *H = Hv.data();
      const double *l1 = l.l1.data(), *l11 = l.l11.data(), *l_1 = l.l_1.data(), *l_11 = l.l_11.data();
      Times(H, l_1 + 1, gz2, M - 1);
      AddTimes(H, l_11 + 1, gz0 + 1, M - 1);
      YplusisCtimesX(gx1 + 1, H, P[1], M - 1);
      YplusisCtimesX(gx2 + 1, H, P[0], M - 1);
      YplusisCtimesX(gx0, gz1, 4 * P[1], M);
      YplusisCtimesX(gx1, gz1, 2 * P[1] + P[0], M);
      YplusisCtimesX(gx2, gz1, 4 * P[1], M);
      Times(H + 1, l1, gz0 + 1, M - 1);
      AddTimes(H + 1, l11, gz2, M - 1);
      YplusisCtimesX(gx0, H + 1, P[0], M - 1);
      YplusisCtimesX(gx1, H + 1, P[1], M - 1);
    }
    for (int k = 0; k < 3; k++) Times(gs + k * M, gs + k * M, G1, M);
  }

  void side(double *Xs, double *X) const {
    const int M = l.M, JX = l.JX;
    Zero(Xs, M);
    if (l.gradients == 1 && l.fjc > 1) {
      refined_taps(Xs, X);
    } else if (l.gradients == 2 && l.fjc > 1) {
      if (l.geometry == GEOM_PLANAR) set_bounds(X);
      refined2(Xs, X);
    } else if (l.gradients == 1) {
      if (l.geometry == GEOM_PLANAR) {
        const double lam = l.lambda;
        YplusisCtimesX(Xs + 1, X, lam, M - 1);
        YplusisCtimesX(Xs, X + 1, lam, M - 1);
        YplusisCtimesX(Xs, X, 1.0 - 2.0 * lam, M);
      } else {
        AddTimes(Xs, X, l.lambda0.data(), M);
        AddTimes(Xs + 1, X, l.lambda_1.data() + 1, M - 1);
        AddTimes(Xs, X + 1, l.lambda1.data(), M - 1);
      }
    } else if (l.gradients == 2) {
      if (l.geometry == GEOM_PLANAR) {
        set_bounds(X);
        if (l.stencil_full) {
          if (l.lattice_type == LAT_SIMPLE_CUBIC) {
            const double C1 = 16.0 / 36.0, C2 = 4.0 / 36.0, C3 = 1.0 / 36.0;
            YplusisCtimesX(Xs, X, C1, M);
            YplusisCtimesX(Xs + 1, X, C2, M - 1);
            YplusisCtimesX(Xs, X + 1, C2, M - 1);
            YplusisCtimesX(Xs + JX, X, C2, M - JX);
            YplusisCtimesX(Xs, X + JX, C2, M - JX);
            YplusisCtimesX(Xs + JX + 1, X, C3, M - JX - 1);
            YplusisCtimesX(Xs + JX, X + 1, C3, M - JX);
            YplusisCtimesX(Xs + 1, X + JX, C3, M - JX);
            YplusisCtimesX(Xs, X + JX + 1, C3, M - JX - 1);
          } else {
            const double Two = 2.0, C = 1.0 / 16.0;
            const int JY = l.JY;
            YplusisCtimesX(Xs, X, Two, M);
            Add(Xs + JX, X, M - JX);
            Add(Xs, X + JX, M - JX);
            Add(Xs + JY, X, M - JY);
            Add(Xs, X + JY, M - JY);
            Norm(Xs, Two, M);
            Add(Xs + JX + JY, X, M - JX - JY);
            Add(Xs, X + JX + JY, M - JX - JY);
            Add(Xs + JX, X + JY, M - JX - JY);
            Add(Xs + JY, X + JX, M - JX - JY);
            Norm(Xs, C, M);
          }
        } else {
          Add(Xs + JX, X, M - JX);
          Add(Xs, X + JX, M - JX);
          Add(Xs + 1, X, M - 1);
          Add(Xs, X + 1, M - 1);
          Norm(Xs, 1.0 / 2.0, M);
          Add(Xs, X, M);
          Norm(Xs, 1.0 / 3.0, M);
        }
      } else {
        const bool sc = l.lattice_type == LAT_SIMPLE_CUBIC;
        const double C1 = sc ? 4.0 / 6.0 : 0.5, C2 = sc ? 1.0 / 6.0 : 0.25, C3 = sc ? 4.0 : 2.0;
        const double *lm = l.lambda_1.data(), *lp = l.lambda1.data();
        YplusisCtimesX(Xs, X, C1, M);
        AddTimes(Xs + JX, X, lm + JX, M - JX);
        AddTimes(Xs, X + JX, lp, M - JX);
        YplusisCtimesX(Xs + 1, X, C2, M - 1);
        YplusisCtimesX(Xs, X + 1, C2, M - 1);
        Norm(Xs, C3, M);
        AddTimes(Xs + JX + 1, X, lm + JX + 1, M - JX - 1);
        AddTimes(Xs + JX, X + 1, lm + JX, M - JX - 1);
        AddTimes(Xs + 1, X + JX, lp + 1, M - JX - 1);
        AddTimes(Xs, X + JX + 1, lp, M - JX - 1);
        Norm(Xs, C2, M);
      }
    } else if (l.fjc > 1) {
      refined3(Xs, X);
    } else if (!(l.stencil_full && l.lattice_type == LAT_SIMPLE_CUBIC)) {
      wide3(Xs, X);
    } else {
      const int JY = l.JY;
      Add(Xs + JX, X, M - JX);
      Add(Xs, X + JX, M - JX);
      Add(Xs + JY, X, M - JY);
      Add(Xs, X + JY, M - JY);
      Add(Xs + 1, X, M - 1);
      Add(Xs, X + 1, M - 1);
      Norm(Xs, 1.0 / 6.0, M);
    }
  }
};

// ------------------------------------------------------------ SVD solve
double pythag(double a, double b) {
  double at = std::fabs(a), bt = std::fabs(b), ct;
  if (at > bt) {
    ct = bt / at;
    return at * std::sqrt(1.0 + ct * ct);
  }
  if (bt > 0.0) {
    ct = at / bt;
    return bt * std::sqrt(1.0 + ct * ct);
  }
  return 0.0;
}
inline double sgn(double a, double b) { return b >= 0.0 ? std::fabs(a) : -std::fabs(a); }

// Golub-Reinsch SVD of an n x n matrix held transposed in a[col][row]; on
// exit a holds U^T, w the singular values, v the right vectors.
bool svd_transposed(std::vector<vec> &a, int n, vec &w, std::vector<vec> &v) {
  const int m = n;
  int flag, i, its, j, jj, k, l = 0, nm = 0;
  double c, f, h, s, x, y, z, anorm = 0.0, g = 0.0, scale = 0.0;
  vec rv1(n, 0.0);
  for (i = 0; i < n; i++) {
    l = i + 1;
    rv1[i] = scale * g;
    g = s = scale = 0.0;
    if (i < m) {
      for (k = i; k < m; k++) scale += std::fabs(a[i][k]);
      if (scale != 0.0) {
        for (k = i; k < m; k++) {
          a[i][k] = a[i][k] / scale;
          s += a[i][k] * a[i][k];
        }
        f = a[i][i];
        g = -sgn(std::sqrt(s), f);
        h = f * g - s;
        a[i][i] = f - g;
        if (i != n - 1) {
          for (j = l; j < n; j++) {
            for (s = 0.0, k = i; k < m; k++) s += a[i][k] * a[j][k];
            f = s / h;
            for (k = i; k < m; k++) a[j][k] += f * a[i][k];
          }
        }
        for (k = i; k < m; k++) a[i][k] = a[i][k] * scale;
      }
    }
    w[i] = scale * g;
    g = s = scale = 0.0;
    if (i < m && i != n - 1) {
      for (k = l; k < n; k++) scale += std::fabs(a[k][i]);
      if (scale != 0.0) {
        for (k = l; k < n; k++) {
          a[k][i] = a[k][i] / scale;
          s += a[k][i] * a[k][i];
        }
        f = a[l][i];
        g = -sgn(std::sqrt(s), f);
        h = f * g - s;
        a[l][i] = f - g;
        for (k = l; k < n; k++) rv1[k] = a[k][i] / h;
        if (i != m - 1) {
          for (j = l; j < m; j++) {
            for (s = 0.0, k = l; k < n; k++) s += a[k][j] * a[k][i];
            for (k = l; k < n; k++) a[k][j] += s * rv1[k];
          }
        }
        for (k = l; k < n; k++) a[k][i] = a[k][i] * scale;
      }
    }
    anorm = std::max(anorm, std::fabs(w[i]) + std::fabs(rv1[i]));
  }
  for (i = n - 1; i >= 0; i--) {
    if (i < n - 1) {
      if (g != 0.0) {
        for (j = l; j < n; j++) v[j][i] = (a[j][i] / a[l][i]) / g;
        for (j = l; j < n; j++) {
          for (s = 0.0, k = l; k < n; k++) s += a[k][i] * v[k][j];
          for (k = l; k < n; k++) v[k][j] += s * v[k][i];
        }
      }
      for (j = l; j < n; j++) v[i][j] = v[j][i] = 0.0;
    }
    v[i][i] = 1.0;
    g = rv1[i];
    l = i;
  }
  for (i = n - 1; i >= 0; i--) {
    l = i + 1;
    g = w[i];
    if (i < n - 1)
      for (j = l; j < n; j++) a[j][i] = 0.0;
    if (g != 0.0) {
      g = 1.0 / g;
      if (i != n - 1) {
        for (j = l; j < n; j++) {
          for (s = 0.0, k = l; k < m; k++) s += a[i][k] * a[j][k];
          f = (s / a[i][i]) * g;
          for (k = i; k < m; k++) a[j][k] += f * a[i][k];
        }
      }
      for (j = i; j < m; j++) a[i][j] = a[i][j] * g;
    } else {
      for (j = i; j < m; j++) a[i][j] = 0.0;
    }
    ++a[i][i];
  }
  for (i = 0; i < n; i++)
    for (j = 0; j < n; j++)
      if (a[i][j] != a[i][j]) return false;
  for (k = n - 1; k >= 0; k--) {
    for (its = 0; its < 30; its++) {
      flag = 1;
      for (l = k; l >= 0; l--) {
        nm = l - 1;
        if (std::fabs(rv1[l]) + anorm == anorm) {
          flag = 0;
          break;
        }
        if (std::fabs(w[nm]) + anorm == anorm) break;
      }
      if (flag) {
        c = 0.0;
        s = 1.0;
        for (i = l; i <= k; i++) {
          f = s * rv1[i];
          if (std::fabs(f) + anorm != anorm) {
            g = w[i];
            h = pythag(f, g);
            w[i] = h;
            h = 1.0 / h;
            c = g * h;
            s = (-f * h);
            for (j = 0; j < m; j++) {
              y = a[nm][j];
              z = a[i][j];
              a[nm][j] = y * c + z * s;
              a[i][j] = z * c - y * s;
            }
          }
        }
      }
      z = w[k];
      if (l == k) {
        if (z < 0.0) {
          w[k] = -z;
          for (j = 0; j < n; j++) v[j][k] = -v[j][k];
        }
        break;
      }
      if (its >= 30) return true;
      x = w[l];
      nm = k - 1;
      y = w[nm];
      g = rv1[nm];
      h = rv1[k];
      f = ((y - z) * (y + z) + (g - h) * (g + h)) / (2.0 * h * y);
      g = pythag(f, 1.0);
      f = ((x - z) * (x + z) + h * ((y / (f + sgn(g, f))) - h)) / x;
      c = s = 1.0;
      for (j = l; j <= nm; j++) {
        i = j + 1;
        g = rv1[i];
        y = w[i];
        h = s * g;
        g = c * g;
        z = pythag(f, h);
        rv1[j] = z;
        c = f / z;
        s = h / z;
        f = x * c + g * s;
        g = g * c - x * s;
        h = y * s;
        y = y * c;
        for (jj = 0; jj < n; jj++) {
          x = v[jj][j];
          z = v[jj][i];
          v[jj][j] = x * c + z * s;
          v[jj][i] = z * c - x * s;
        }
        z = pythag(f, h);
        w[j] = z;
        if (z != 0.0) {
          z = 1.0 / z;
          c = f * z;
          s = h * z;
        }
        f = (c * g) + (s * y);
        x = (c * y) - (s * g);
        for (jj = 0; jj < m; jj++) {
          y = a[j][jj];
          z = a[i][jj];
          a[j][jj] = y * c + z * s;
          a[i][jj] = z * c - y * s;
        }
      }
      rv1[l] = 0.0;
      rv1[k] = f;
      w[k] = x;
    }
  }
  return true;
}

// Solve A c = 1 through the SVD (SFNewton::Ax src/sfnewton.cpp:793-835).
// Returns false on a NaN (the reference throws -2 / -3).
bool solve_ones(const vec &A, int N, vec &X) {
  for (int i = 0; i < N * N; i++)
    if (A[i] != A[i]) return false;
  X.assign(N, 0.0);
  if (N == 1) {
    X[0] = 1;
    return true;
  }
  std::vector<vec> U(N, vec(N)), V(N, vec(N, 0.0));
  vec S(N, 0.0);
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++) U[i][j] = A[i * N + j];
  if (!svd_transposed(U, N, S, V)) return false;
  for (int i = 0; i < N; i++) {
    X[i] = 0;
    for (int j = 0; j < N; j++) X[i] += U[i][j];
    S[i] = X[i] / S[i];
    X[i] = 0;
  }
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++) X[i] += V[i][j] * S[j];
  return true;
}

// ----------------------------------------------------------------- system
struct Oracle {
  SystemSpec s;
  int M = 0, nseg = 0, nmol = 0, nit = 0;
  std::vector<vec> u, G1, phi, phi_side;  // per segment
  vec seg_phibulk;
  std::vector<vec> mol_phi;               // per mol: molmon * M
  std::vector<vec> mol_phitot;
  vec phitot, alpha, F_dens, GP_dens;
  vec psi, EE, q, eps, Xel;               // electrostatics (charged systems)
  vec GN, norm, n, theta, phibulk, Mu, theta_out;
  double F = 0, GP = 0, residual = 0;
  int status = 0;  // 0 ok, -1 GN<=0, -2/-3 NaN in DIIS, -4 negative solvent phibulk
  std::vector<vec> Gf;                    // scratch per chain
  vec Gb;

  explicit Oracle(const SystemSpec &spec) : s(spec) {
    M = s.lat.M;
    nseg = (int)s.seg.size();
    nmol = (int)s.mol.size();
    nit = (int)s.itmon.size();
    u.assign(nseg, vec(M, 0.0));
    G1.assign(nseg, vec(M, 0.0));
    phi.assign(nseg, vec(M, 0.0));
    phi_side.assign(nseg, vec(M, 0.0));
    seg_phibulk.assign(nseg, 0.0);
    mol_phi.resize(nmol);
    mol_phitot.assign(nmol, vec(M, 0.0));
    for (int i = 0; i < nmol; i++) mol_phi[i].assign(s.mol[i].molmon.size() * M, 0.0);
    phitot.assign(M, 0.0);
    psi.assign(M, 0.0);
    EE.assign(M, 0.0);
    q.assign(M, 0.0);
    eps.assign(M, 0.0);
    Xel.assign(M, 0.0);
    alpha.assign(M, 0.0);
    F_dens.assign(M, 0.0);
    GP_dens.assign(M, 0.0);
    GN.assign(nmol, 0.0);
    norm.assign(nmol, 0.0);
    n.assign(nmol, 0.0);
    theta.assign(nmol, 0.0);
    phibulk.assign(nmol, 0.0);
    Mu.assign(nmol, 0.0);
    theta_out.assign(nmol, 0.0);
    for (int i = 0; i < nmol; i++) {
      n[i] = s.mol[i].n;
      theta[i] = s.mol[i].theta;
      phibulk[i] = s.mol[i].phibulk;
    }
  }

  // Molecule::fraction (src/molecule.cpp:2686-2697); dendrimer blocks count
  // their degeneracy (mol_dend::fraction src/mol_dendrimer.cpp:12-22)
  double fraction(int mol, int seg) const {
    const MoleculeSpec &m = s.mol[mol];
    int Nseg = 0;  // a ring's first segment is not counted (src/molecule.cpp:2686-2697)
    for (size_t b = m.ring ? 1 : 0; b < m.blocks.size(); b++)
      if (m.blocks[b].seg == seg) Nseg += m.blocks[b].n * (b < m.block_deg.size() ? m.block_deg[b] : 1);
    return 1.0 * Nseg / m.chainlength;
  }

  // mol_branched::ComputePhi (src/mol_branched.cpp:303-420),
  // mol_dend::BackAndForth (src/mol_dendrimer.cpp:122-188) and
  // mol_comb::GoBackAndForth (src/mol_comb.cpp:206-400) by the row
  // schedule of the molecule (build_branch_plan: the reference's recursion
  // unrolled into row operations); Markov 1.  UNITY is a row of ones
  // (the reference's UNITY array), Initiate / Terminate are copies.
  void compute_phi_branched(int mi) {
    const MoleculeSpec &m = s.mol[mi];
    Lat lat(s.lat);
    vec &rho = mol_phi[mi];
    std::fill(rho.begin(), rho.end(), 0.0);
    BranchPlan plan;
    std::string e;
    if (!build_branch_plan(m, plan, e)) {
      status = 3;
      return;
    }
    std::vector<vec> R(plan.rows, vec(M, 0.0));
    const vec unity(M, 1.0);
    auto run = [&](const std::vector<BranchOp> &ops) {
      for (const BranchOp &o : ops) {
        switch (o.kind) {
          case BR_INIT: Cp(R[o.a].data(), G1[o.seg].data(), M); break;
          case BR_PROP: lat.propagate(R[o.a].data(), R[o.b].data(), o.seg < 0 ? unity.data() : G1[o.seg].data()); break;
          case BR_COPY: Cp(R[o.a].data(), R[o.b].data(), M); break;
          case BR_TIMES: Times(R[o.a].data(), R[o.b].data(), R[o.c].data(), M); break;
          case BR_ONE: Cp(R[o.a].data(), unity.data(), M); break;
          case BR_TIMES_G1: Times(R[o.a].data(), R[o.b].data(), G1[o.seg].data(), M); break;
          case BR_DIV: Div(R[o.a].data(), R[o.b].data(), M); break;
          case BR_DIV_G1: Div(R[o.a].data(), G1[o.seg].data(), M); break;
          default:
            // AddTimes (f = 1) or YplusisCtimesAtimesB (a degeneracy): rho += (f Gf) Gb
            if (o.f == 1.0)
              AddTimes(rho.data() + o.c * M, R[o.a].data(), R[o.b].data(), M);
            else
              YplusisCtimesAtimesB(rho.data() + o.c * M, R[o.a].data(), R[o.b].data(), o.f, M);
            break;
        }
      }
    };
    run(plan.forward);
    GN[mi] = lat.weighted_sum(R[plan.gn_row].data());
    run(plan.backward);
  }

  // Molecule::ComputePhi for a linear chain; accumulates Gf*Gb into mol_phi
  // mol_linear::ComputePhi for a ring (src/mol_linear.cpp:35-111), first
  // order: for every interior site x where the first segment
  // may sit, a chain starts there (G0 = delta_x G1), runs through the middle
  // blocks and ends on x with the last segment (the row masked to x); GN is
  // the sum over x, the merged end segment's density is G0 times that row,
  // and the middle segments' densities come from the backward pass started
  // at G0.  Rows: 0 .. N (N = chainlength, one less than the composition).
  void compute_phi_ring(int mi) {
    const MoleculeSpec &m = s.mol[mi];
    Lat lat(s.lat);
    const int N = m.chainlength;
    vec &rho = mol_phi[mi];
    std::fill(rho.begin(), rho.end(), 0.0);
    std::vector<int> seg_of, mm_of;
    for (size_t b = 0; b < m.blocks.size(); b++)
      for (int j = 0; j < m.blocks[b].n; j++) {
        seg_of.push_back(m.blocks[b].seg);
        mm_of.push_back(m.block_molmon[b]);
      }
    if ((int)Gf.size() < N + 1) Gf.resize(N + 1);
    for (int k = 0; k <= N; k++) Gf[k].assign(M, 0.0);
    Gb.assign(2 * M, 0.0);
    vec G0(M, 0.0);
    const double *g1a = G1[seg_of[0]].data();
    double gn = 0;
    // interior sites in the reference's order: z outer, x inner
    std::vector<int> starts;
    const LatticeSpec &L = s.lat;
    const int nz = L.gradients == 3 ? L.MZ : 1, ny = L.gradients >= 2 ? L.MY : 1;
    for (int z = 0; z < nz; z++)
      for (int y = 0; y < ny; y++)
        for (int x = 1; x < L.MX + 1; x++) {
          if (L.gradients == 1) starts.push_back(x);
          else if (L.gradients == 2) starts.push_back(x * L.JX + (y + 1) * L.JY);
          else starts.push_back(x * L.JX + (y + 1) * L.JY + (z + 1));
        }
    for (int x : starts) {
      if (!(g1a[x] > 0)) continue;
      Zero(G0.data(), M);
      G0[x] = g1a[x];
      Cp(Gf[0].data(), G0.data(), M);
      for (int k = 1; k <= N; k++) lat.propagate(Gf[k].data(), Gf[k - 1].data(), G1[seg_of[k]].data());
      double *last = Gf[N].data();
      for (int z = 0; z < M; z++)
        if (z != x) last[z] = 0.0;
      gn += lat.weighted_sum(last);
      double *gb = Gb.data() + (N % 2) * M;
      Cp(gb, G0.data(), M);
      YplusisCtimesAtimesB(rho.data() + mm_of[0] * M, gb, last, 1.0, M);
      for (int k = N - 1; k >= 1; k--) {
        gb = Gb.data() + (k % 2) * M;
        lat.propagate(gb, Gb.data() + ((k + 1) % 2) * M, G1[seg_of[k]].data());
        YplusisCtimesAtimesB(rho.data() + mm_of[k] * M, Gf[k].data(), gb, 1.0, M);
      }
    }
    GN[mi] = gn;
  }

  void compute_phi_mol(int mi) {
    const MoleculeSpec &m = s.mol[mi];
    if (m.scheduled()) {
      compute_phi_branched(mi);
      return;
    }
    if (m.ring) {  // its blocks hold one segment more than the chain
      compute_phi_ring(mi);
      return;
    }
    Lat lat(s.lat);
    const int N = m.chainlength, size = m.size;
    const int W = size * M;
    vec &rho = mol_phi[mi];
    std::fill(rho.begin(), rho.end(), 0.0);
    // segment index per s
    std::vector<int> seg_of(N), mm_of(N);
    {
      int k = 0;
      for (size_t b = 0; b < m.blocks.size(); b++)
        for (int j = 0; j < m.blocks[b].n; j++) {
          seg_of[k] = m.blocks[b].seg;
          mm_of[k] = m.block_molmon[b];
          k++;
        }
    }
    if ((int)Gf.size() < N) Gf.resize(N);
    for (int k = 0; k < N; k++) Gf[k].assign(W, 0.0);
    Gb.assign(2 * W, 0.0);
    // forward
    for (int k = 0; k < N; k++) {
      const double *g1 = G1[seg_of[k]].data();
      if (k == 0) {
        for (int c = 0; c < size; c++) Cp(Gf[0].data() + c * M, g1, M);
      } else if (m.markov == 2) {
        lat.propagateF(Gf[k].data(), Gf[k - 1].data(), g1, m.P.data());
      } else {
        lat.propagate(Gf[k].data(), Gf[k - 1].data(), g1);
      }
    }
    // GN (LGrad1::ComputeGN src/LGrad1.cpp:943-955 and analogues)
    double *Glast = Gf[N - 1].data();
    if (m.markov == 2 && s.lat.gradients == 3) {
      // LGrad3::ComputeGN (src/LGrad3.cpp:1883-1895): the mean of the 6 components
      double g = 0;
      for (int k = 0; k < 6; k++) g += lat.weighted_sum(Glast + k * M);
      GN[mi] = g / (1.0 * 6);
    } else if (m.markov == 2 && s.lat.gradients == 2) {
      // LGrad2::ComputeGN (src/LGrad2.cpp:1814-1840): (G0 + G1 + 2 G2 + G3 + G4) / 6
      double g = 0;
      for (int k = 0; k < 5; k++) g += (k == 2 ? 2 * lat.weighted_sum(Glast + k * M) : lat.weighted_sum(Glast + k * M));
      GN[mi] = g / 6.0;
    } else if (m.markov == 2 && s.lat.lattice_type == LAT_HEXAGONAL) {
      const int F = s.lat.FJC;
      double g = lat.weighted_sum(Glast);
      for (int k = 1; k < F - 1; k++) g += 2.0 * lat.weighted_sum(Glast + k * M);
      g += lat.weighted_sum(Glast + (F - 1) * M);
      GN[mi] = g / (4.0 * s.lat.fjc);
    } else if (m.markov == 2) {
      double g = lat.weighted_sum(Glast);
      g += 4.0 * lat.weighted_sum(Glast + M);
      g += lat.weighted_sum(Glast + 2 * M);
      GN[mi] = g / 6.0;
    } else {
      GN[mi] = lat.weighted_sum(Glast);
    }
    // backward
    for (int k = N - 1; k >= 0; k--) {
      double *gb = Gb.data() + (k % 2) * W;
      const double *g1 = G1[seg_of[k]].data();
      if (k == N - 1) {
        for (int c = 0; c < size; c++) Cp(gb + c * M, g1, M);
      } else if (m.markov == 2) {
        lat.propagateB(gb, Gb.data() + ((k + 1) % 2) * W, g1, m.P.data());
      } else {
        lat.propagate(gb, Gb.data() + ((k + 1) % 2) * W, g1);
      }
      double *r = rho.data() + mm_of[k] * M;
      const double *gf = Gf[k].data();
      if (m.markov == 2 && s.lat.gradients == 3) {
        // LGrad3::AddPhiS (src/LGrad3.cpp:1897-1905)
        for (int k = 0; k < 6; k++) YplusisCtimesAtimesB(r, gf + k * M, gb + k * M, 1.0 / 6.0, M);
      } else if (m.markov == 2 && s.lat.gradients == 2) {
        // LGrad2::AddPhiS (src/LGrad2.cpp:1849-1882)
        for (int k = 0; k < 5; k++)
          YplusisCtimesAtimesB(r, gf + k * M, gb + k * M, k == 2 ? 1.0 / 3.0 : 1.0 / 6.0, M);
      } else if (m.markov == 2 && s.lat.lattice_type == LAT_HEXAGONAL) {
        // LGrad1::AddPhiS (src/LGrad1.cpp:876-898)
        const int F = s.lat.FJC;
        const double C1 = s.lat.fjc == 1 ? 1.0 / 4.0 : 0.5 / (F - 1.0);
        const double C2 = s.lat.fjc == 1 ? 2.0 / 4.0 : 1.0 / (F - 1.0);
        YplusisCtimesAtimesB(r, gf, gb, C1, M);
        for (int k = 1; k < F - 1; k++) YplusisCtimesAtimesB(r, gf + k * M, gb + k * M, C2, M);
        YplusisCtimesAtimesB(r, gf + (F - 1) * M, gb + (F - 1) * M, C1, M);
      } else if (m.markov == 2) {
        YplusisCtimesAtimesB(r, gf, gb, 1.0 / 6.0, M);
        YplusisCtimesAtimesB(r, gf + M, gb + M, 4.0 / 6.0, M);
        YplusisCtimesAtimesB(r, gf + 2 * M, gb + 2 * M, 1.0 / 6.0, M);
      } else {
        AddTimes(r, gf, gb, M);
      }
    }
  }

  // ------------------------------------------------------ electrostatics
  // Physical constants of the reference (src/namics.cpp:43-49); the bond
  // length is the lattice's `bondlength` over fjc (src/lattice.cpp:332-337).
  static constexpr double kE = 1.60217e-19, kT = 298.15, kB = 1.38065e-23, kEps0 = 8.85418e-12;
  double bond() const { return s.lat.bond_length / s.lat.fjc; }

  // LGrad1::set_M_bounds (src/LGrad1.cpp:820-832): mirrored ghosts whatever
  // the boundary conditions
  // LGrad2::set_M_bounds (src/LGrad2.cpp:1702-1723) and LGrad3::set_M_bounds
  // (src/LGrad3.cpp:1679-1702) with fjc = 1: faces mirrored; in 2-D the
  // corners from the x-ghost rows, in 3-D the edges left as they are.
  void set_M_bounds(double *X) const {
    const LatticeSpec &l = s.lat;
    const int f = l.fjc, MX = l.MX, MY = l.MY, MZ = l.MZ, JX = l.JX, JY = l.JY;
    if (l.gradients == 2) {
      for (int x = 1; x < MX + 1; x++) {
        X[x * JX + 0] = X[x * JX + 1];
        X[x * JX + MY + 1] = X[x * JX + MY];
      }
      for (int y = 1; y < MY + 1; y++) {
        X[0 + y] = X[1 * JX + y];
        X[(MX + 1) * JX + y] = X[MX * JX + y];
      }
      X[0] = X[1];
      X[MY + 1] = X[MY];
      X[(MX + 1) * JX + 0] = X[(MX + 1) * JX + 1];
      X[(MX + 1) * JX + MY + 1] = X[(MX + 1) * JX + MY];
      return;
    }
    if (l.gradients == 3) {
      for (int x = 1; x < MX + 1; x++)
        for (int y = 1; y < MY + 1; y++) {
          X[x * JX + y * JY + 0] = X[x * JX + y * JY + 1];
          X[x * JX + y * JY + MZ + 1] = X[x * JX + y * JY + MZ];
        }
      for (int y = 1; y < MY + 1; y++)
        for (int z = 1; z < MZ + 1; z++) {
          X[0 + y * JY + z] = X[1 * JX + y * JY + z];
          X[(MX + 1) * JX + y * JY + z] = X[MX * JX + y * JY + z];
        }
      for (int z = 1; z < MZ + 1; z++)
        for (int x = 1; x < MX + 1; x++) {
          X[x * JX + 0 + z] = X[x * JX + 1 * JY + z];
          X[x * JX + (MY + 1) * JY + z] = X[x * JX + MY * JY + z];
        }
      return;
    }
    for (int k = 0; k < f; k++) {
      X[k] = X[2 * f - k - 1];
      X[MX + f + k] = X[MX + f - k - 1];
    }
  }

  // LG1Planar::UpdateEE (src/LG1Planar.cpp:162-178) and LGrad1::UpdateEE
  // (src/LGrad1.cpp:592-651): the field energy density of psi at each site.
  void update_EE() {
    const int f = s.lat.fjc, MX = s.lat.MX;
    const double k_BT = kB * kT;
    double pf = 0.5 * kEps0 * bond() / k_BT * (k_BT / kE) * (k_BT / kE);
    double *P = psi.data();
    set_M_bounds(P);
    Zero(EE.data(), M);
    const int geo = s.lat.geometry;
    const double off = s.lat.offset_first_layer;  // offset_first_layer * fjc
    const int MY = s.lat.MY, JX = s.lat.JX, JY = s.lat.JY;
    if (s.lat.gradients == 3) {
      // LGrad3::UpdateEE (src/LGrad3.cpp:1131-1141): AddGradSquare along
      // each axis over the flattened array, then the prefactor
      for (int i = 0; i < M - 2; i++) EE[i + 1] += std::pow(P[i] - P[i + 1], 2) + std::pow(P[i + 1] - P[i + 2], 2);
      for (int i = 0; i < M - 2 * JX; i++)
        EE[i + JX] += std::pow(P[i] - P[i + JX], 2) + std::pow(P[i + JX] - P[i + 2 * JX], 2);
      for (int i = 0; i < M - 2 * JY; i++)
        EE[i + JY] += std::pow(P[i] - P[i + JY], 2) + std::pow(P[i + JY] - P[i + 2 * JY], 2);
      Norm(EE.data(), pf, M);
      return;
    }
    if (s.lat.gradients == 2 && geo == GEOM_PLANAR) {
      // LG2Planar::UpdateEE (src/LG2Planar.cpp:1130-1152)
      pf = pf / 2.0;
      for (int x = f; x < MX + f; x++)
        for (int y = f; y < MY + f; y++) {
          const int z = x * JX + y;
          double Exmin = P[z] - P[z - JX];
          Exmin *= Exmin;
          double Explus = P[z] - P[z + JX];
          Explus *= Explus;
          double Eymin = P[z] - P[z - 1];
          Eymin *= Eymin;
          double Eyplus = P[z] - P[z + 1];
          Eyplus *= Eyplus;
          EE[z] = pf * (Exmin + Explus + Eymin + Eyplus);
        }
      return;
    }
    if (s.lat.gradients == 2) {
      // LGrad2::UpdateEE (src/LGrad2.cpp:1392-1417), integer radius
      pf = pf / 2;
      int r = (int)off;
      for (int x = f; x < MX + f; x++) {
        r++;
        for (int y = f; y < MY + f; y++) {
          const int z = x * JX + y;
          double Exmin = P[z] - P[z - JX];
          Exmin *= (r - 0.5) * Exmin * 2 * kPIE;
          double Explus = P[z] - P[z + JX];
          Explus *= (r + 0.5) * Explus * 2 * kPIE;
          double Eymin = (P[z] - P[z - 1]) * f;
          Eymin *= Eymin;
          double Eyplus = (P[z] - P[z + 1]) * f;
          Eyplus *= Eyplus;
          EE[z] = pf * ((Exmin + Explus) / s.lat.L[z] + Eymin + Eyplus);
        }
      }
      return;
    }
    if (geo == GEOM_PLANAR) {
      pf = pf / 2.0 * f * f;
      double Explus = P[f - 1] - P[f];
      Explus *= Explus;
      for (int x = f; x < MX + f; x++) {
        double Exmin = Explus;
        Explus = P[x] - P[x + 1];
        Explus *= Explus;
        EE[x] = pf * (Exmin + Explus);
      }
    } else if (geo == GEOM_CYLINDRICAL) {
      int r = (int)off;  // an int in the reference
      pf = pf * kPIE;
      for (int x = f; x < MX + f; x++) {
        r++;
        double Exmin = P[x] - P[x - 1];
        Exmin *= (r - 1) * Exmin;
        double Explus = P[x] - P[x + 1];
        Explus *= r * Explus;
        EE[x] = pf * (Exmin + Explus) / s.lat.L[x];
      }
    } else {
      pf = pf * kPIE * 2 / f;
      int r = (int)(off + 1.0);  // an int in the reference
      double Explus = r * (P[f] - P[f + 1]);
      Explus *= Explus;
      EE[f] = pf * Explus / s.lat.L[f];
      for (int x = f + 1; x < MX + f; x++) {
        r += 1;
        double Exmin = Explus;
        Explus = r * (P[x] - P[x + 1]);
        Explus *= Explus;
        EE[x] = pf * (Exmin + Explus) / s.lat.L[x];
      }
    }
  }

  // System::DoElectrostatics (src/system.cpp:1920-1953) without fixed
  // potentials: charge density and dielectric constant per site, psi from x.
  void do_electrostatics(double *gp, const double *xp) {
    Lat lat(s.lat);
    Zero(q.data(), M);
    Zero(eps.data(), M);
    for (int i = 0; i < nseg; i++) {
      YplusisCtimesX(q.data(), phi[i].data(), s.seg[i].valence, M);
      lat.set_bounds(phi[i].data());
      YplusisCtimesX(eps.data(), phi[i].data(), s.seg[i].epsilon, M);
    }
    Div(q.data(), phitot.data(), M);
    Div(eps.data(), phitot.data(), M);
    Cp(psi.data(), xp, M);
    Cp(gp, psi.data(), M);
    set_M_bounds(psi.data());
    // fixed surface potentials: Segment::UpdateValence (src/segment.cpp:1972-1979)
    // for every frozen segment in order
    if (s.fixed_psi0)
      for (int i = 0; i < nseg; i++)
        if (s.seg[i].freedom == SEG_FROZEN && s.seg[i].fixed_psi0) {
          // OverwriteC (src/tools_host.h:514-517): psi0 on the mask, zero elsewhere
          for (int z = 0; z < M; z++) psi[z] = s.seg[i].mask[z] == 1 ? s.seg[i].psi0 : 0.0;
          set_M_bounds(psi.data());
          update_q(gp, s.seg[i].mask);
        }
  }

  // LG1Planar::UpdateQ (src/LG1Planar.cpp:226-244), LGrad1::UpdateQ
  // (src/LGrad1.cpp:733-767): the charge on the mask that the fixed
  // potential needs, and g = -q there.  The spherical correction is never
  // applied (its geometry test misspells "spherical").
  void update_q(double *gp, const std::vector<int> &mask) {
    const int f = s.lat.fjc, MX = s.lat.MX;
    const double k_BT = kB * kT;
    const double C = -kE * kE / (kEps0 * k_BT * bond());
    const double *P = psi.data(), *E = eps.data();
    double a, b, c;
    const int MY = s.lat.MY, MZ = s.lat.MZ, JX = s.lat.JX, JY = s.lat.JY;
    if (s.lat.gradients == 3) {
      // LGrad3::UpdateQ (src/LGrad3.cpp:1314-1347): the loops stop one layer
      // short of each upper face, the charge on the mask is written into psi
      // (not q), and g = -q with the q of DoElectrostatics
      double *Q = psi.data();
      for (int x = 1; x < MX; x++)
        for (int y = 1; y < MY; y++) {
          double epsZplus = E[x * JX + y * JY] + E[x * JX + y * JY + 1];
          for (int z = 1; z < MZ; z++) {
            const int i = x * JX + y * JY + z;
            const double epsZmin = epsZplus;
            epsZplus = E[i] + E[i + 1];
            const double epsYmin = E[i] + E[i - JY], epsYplus = E[i] + E[i + JY];
            const double epsXmin = E[i] + E[i - JX], epsXplus = E[i] + E[i + JX];
            if (mask[i] == 1) {
              Q[i] = (epsXmin * Q[i - JX] + epsXplus * Q[i + JX] + epsYmin * Q[i - JY] + epsYplus * Q[i + JY] +
                      epsZmin * Q[i - 1] + epsZplus * Q[i + 1] -
                      (epsXmin + epsXplus + epsYmin + epsYplus + epsZmin + epsZplus) * Q[i]) /
                     C;
              gp[i] = -q[i];
            }
          }
        }
      return;
    }
    if (s.lat.gradients == 2) {
      // LGrad2::UpdateQ (src/LGrad2.cpp:1447-1484; its radial term reads eps
      // at index x) and LG2Planar::UpdateQ (src/LG2Planar.cpp:1211-1236)
      for (int x = f; x < MX + f; x++)
        for (int y = f; y < MY + f; y++) {
          const int i = x * JX + y;
          if (mask[i] == 1) q[i] = -0.5 * (P[i - JX] + P[i + JX] + P[i - 1] + P[i + 1] - 4 * P[i]) * f * f * E[i] / C;
        }
      if (s.lat.geometry != GEOM_PLANAR)
        for (int x = f; x < MX + f; x++)
          for (int y = f; y < MY + f; y++) {
            const int i = x * JX + y;
            if (mask[i] == 1) q[i] -= (P[i + JX] - P[i - JX]) / (2.0 * (s.lat.offset_first_layer + x - f + 0.5)) * f * E[x] / C;
          }
      if (s.grad_epsilon)
        for (int x = f; x < MX + f; x++)
          for (int y = f; y < MY + f; y++) {
            const int i = x * JX + y;
            if (mask[i] == 1)
              q[i] -= 0.25 * (E[i + JX] - E[i - JX]) * (P[i + JX] - P[i - JX] + E[i + 1] - E[i - 1]) * (P[i + 1] - P[i - 1]) * f * f / C;
          }
      for (int x = f; x < MX + f; x++)
        for (int y = f; y < MY + f; y++) {
          const int i = x * JX + y;
          if (mask[i] == 1) gp[i] = -q[i];
        }
      return;
    }
    b = P[f - 1];
    c = P[f];
    for (int x = f; x < MX + f; x++) {
      a = b;
      b = c;
      c = P[x + 1];
      if (mask[x] == 1) q[x] = -0.5 * (a - 2 * b + c) * f * f * E[x] / C;
    }
    if (s.lat.geometry == GEOM_CYLINDRICAL) {
      b = P[f - 1];
      c = P[f];
      for (int x = f; x < MX + f; x++) {
        a = b;
        b = c;
        c = P[x + 1];
        if (mask[x] == 1) q[x] -= (c - a) / (2.0 * (s.lat.offset_first_layer + x - f + 0.5)) * f * E[x] / C;
      }
    }
    if (s.grad_epsilon) {
      double a_, b_, c_;
      b = P[f - 1];
      c = P[f];
      b_ = E[f - 1];
      c_ = E[f];
      for (int x = f; x < MX + f; x++) {
        a = b;
        b = c;
        c = P[x + 1];
        a_ = b_;
        b_ = c_;
        c_ = E[x + 1];
        if (mask[x] == 1) q[x] -= 0.25 * (c_ - a_) * (c - a) * f * f / C;
      }
    }
    for (int x = f; x < MX + f; x++)
      if (mask[x] == 1) gp[x] = -q[x];
  }

  // UpdatePsi with fixed potentials (src/LG1Planar.cpp:200-222,
  // src/LGrad1.cpp:687-729): Jacobi sweeps of psi off the potential masks
  // (the cylindrical correction only on cylindrical lattices, as above),
  // g -= the new psi there.
  void update_psi_fixed(double *gp, const std::vector<int> &pmask) {
    const int f = s.lat.fjc, MX = s.lat.MX;
    const double k_BT = kB * kT;
    const double C = kE * kE / (kEps0 * k_BT * bond());
    double *P = psi.data();
    const double *E = eps.data();
    double a, b, c;
    const int MY = s.lat.MY, MZ = s.lat.MZ, JX = s.lat.JX, JY = s.lat.JY;
    if (s.lat.gradients == 3) {
      // LGrad3::UpdatePsi with fixed potentials (src/LGrad3.cpp:1239-1282):
      // the symmetric Gauss-Seidel sweep off the masks
      auto site = [&](int i, double epsZmin, double epsZplus) {
        const double epsYmin = E[i] + E[i - JY], epsYplus = E[i] + E[i + JY];
        const double epsXmin = E[i] + E[i - JX], epsXplus = E[i] + E[i + JX];
        P[i] = (epsXmin * P[i - JX] + epsXplus * P[i + JX] + epsYmin * P[i - JY] + epsYplus * P[i + JY] +
                epsZmin * P[i - 1] + epsZplus * P[i + 1] + C * q[i]) /
               (epsXmin + epsXplus + epsYmin + epsYplus + epsZmin + epsZplus);
      };
      for (int x = 1; x < MX + 1; x++)
        for (int y = 1; y < MY + 1; y++) {
          double epsZplus = E[x * JX + y * JY] + E[x * JX + y * JY + 1];
          for (int z = 1; z < MZ + 1; z++) {
            const int i = x * JX + y * JY + z;
            const double epsZmin = epsZplus;
            epsZplus = E[i] + E[i + 1];
            if (pmask[i] == 0) site(i, epsZmin, epsZplus);
          }
        }
      for (int x = MX; x > 0; x--)
        for (int y = MY; y > 0; y--) {
          double epsZmin = E[x * JX + y * JY + MZ + 1] + E[x * JX + y * JY + MZ];
          for (int z = MZ; z > 0; z--) {
            const int i = x * JX + y * JY + z;
            const double epsZplus = epsZmin;
            epsZmin = E[i] + E[i - 1];
            if (pmask[i] == 0) {
              site(i, epsZmin, epsZplus);
              gp[i] -= P[i];
            }
          }
        }
      return;
    }
    if (s.lat.gradients == 2) {
      // LGrad2 / LG2Planar::UpdatePsi with fixed potentials
      // (src/LGrad2.cpp:1448-1482, src/LG2Planar.cpp:1182-1206): Jacobi off
      // the masks, the radial correction on cylindrical lattices
      double *X = Xel.data();
      for (int x = f; x < MX + f; x++)
        for (int y = f; y < MY + f; y++) {
          const int i = x * JX + y;
          if (pmask[i] == 0) X[i] = 0.25 * (P[i - JX] + P[i + JX] + P[i - 1] + P[i + 1]) + 0.5 * q[i] * C / E[i];
        }
      if (s.lat.geometry != GEOM_PLANAR)
        for (int x = f; x < MX + f; x++)
          for (int y = f; y < MY + f; y++) {
            const int i = x * JX + y;
            if (pmask[i] == 0) X[i] += (P[i + JX] - P[i - JX]) / (2.0 * (s.lat.offset_first_layer + x - f + 0.5)) * f;
          }
      if (s.grad_epsilon)
        for (int x = f; x < MX + f; x++)
          for (int y = f; y < MY + f; y++) {
            const int i = x * JX + y;
            if (pmask[i] == 0)
              X[i] += 0.25 * (E[i + JX] - E[i - JX]) * (P[i + JX] - P[i - JX] + E[i + 1] - E[i - 1]) * (P[i + 1] - P[i - 1]) / E[i] * f * f;
          }
      for (int x = f; x < MX + f; x++)
        for (int y = f; y < MY + f; y++) {
          const int i = x * JX + y;
          if (pmask[i] == 0) {
            P[i] = X[i];
            gp[i] -= P[i];
          }
        }
      return;
    }
    b = P[f - 1];
    c = P[f];
    for (int x = f; x < MX + f; x++) {
      a = b;
      b = c;
      c = P[x + 1];
      if (pmask[x] == 0) P[x] = 0.5 * (a + c) + q[x] * C / E[x];
    }
    if (s.lat.geometry == GEOM_CYLINDRICAL) {
      b = P[f - 1];
      c = P[f];
      for (int x = f; x < MX + f; x++) {
        a = b;
        b = c;
        c = P[x + 1];
        if (pmask[x] == 0) P[x] += (c - a) / (2.0 * (s.lat.offset_first_layer + x - f + 0.5)) * f;
      }
    }
    if (s.grad_epsilon) {
      double a_, b_, c_;
      b = P[f - 1];
      c = P[f];
      b_ = E[f - 1];
      c_ = E[f];
      for (int x = f; x < MX + f; x++) {
        a = b;
        b = c;
        c = P[x + 1];
        a_ = b_;
        b_ = c_;
        c_ = E[x + 1];
        if (pmask[x] == 0) P[x] += 0.25 * (c_ - a_) * (c - a) / E[x] * f * f;
      }
    }
    for (int x = f; x < MX + f; x++)
      if (pmask[x] == 0) gp[x] -= P[x];
  }

  // LG1Planar::UpdatePsi (src/LG1Planar.cpp:181-199) and LGrad1::UpdatePsi
  // (src/LGrad1.cpp:654-686) without fixed potentials: one sweep of the
  // discretised Poisson equation (upwind in planar and cylindrical
  // geometry, Jacobi in spherical), g -= the swept psi.
  void update_psi(double *gp) {
    const int f = s.lat.fjc, MX = s.lat.MX;
    const double k_BT = kB * kT;
    double C = kE * kE / (kEps0 * k_BT * bond());
    const double *P = psi.data(), *E = eps.data();
    double *X = Xel.data();
    const int geo = s.lat.geometry;
    const double off = s.lat.offset_first_layer;
    const int MY = s.lat.MY, MZ = s.lat.MZ, JX = s.lat.JX, JY = s.lat.JY;
    if (s.lat.gradients == 3) {
      // LGrad3::UpdatePsi (src/LGrad3.cpp:1150-1224): symmetric Gauss-Seidel,
      // a forward lexicographic sweep and a backward one, in place; g -= psi
      // in the backward sweep
      double *Q = psi.data();
      auto site = [&](int x, int y, int z, double epsZmin, double epsZplus) {
        const int i = x * JX + y * JY + z;
        const double epsYmin = E[i] + E[i - JY], epsYplus = E[i] + E[i + JY];
        const double epsXmin = E[i] + E[i - JX], epsXplus = E[i] + E[i + JX];
        Q[i] = (epsXmin * Q[i - JX] + epsXplus * Q[i + JX] + epsYmin * Q[i - JY] + epsYplus * Q[i + JY] +
                epsZmin * Q[i - 1] + epsZplus * Q[i + 1] + C * q[i]) /
               (epsXmin + epsXplus + epsYmin + epsYplus + epsZmin + epsZplus);
      };
      for (int x = 1; x < MX + 1; x++)
        for (int y = 1; y < MY + 1; y++) {
          double epsZplus = E[x * JX + y * JY] + E[x * JX + y * JY + 1];
          for (int z = 1; z < MZ + 1; z++) {
            const double epsZmin = epsZplus;
            epsZplus = E[x * JX + y * JY + z] + E[x * JX + y * JY + z + 1];
            site(x, y, z, epsZmin, epsZplus);
          }
        }
      for (int x = MX; x > 0; x--)
        for (int y = MY; y > 0; y--) {
          double epsZmin = E[x * JX + y * JY + MZ + 1] + E[x * JX + y * JY + MZ];
          for (int z = MZ; z > 0; z--) {
            const double epsZplus = epsZmin;
            epsZmin = E[x * JX + y * JY + z] + E[x * JX + y * JY + z - 1];
            site(x, y, z, epsZmin, epsZplus);
            gp[x * JX + y * JY + z] -= Q[x * JX + y * JY + z];
          }
        }
      return;
    }
    if (s.lat.gradients == 2 && geo == GEOM_PLANAR) {
      // LG2Planar::UpdatePsi (src/LG2Planar.cpp:1163-1181): upwind in x and y
      C = C * 2.0 / f / f;
      for (int x = f; x < MX + f; x++)
        for (int y = f; y < MY + f; y++) {
          const int i = x * JX + y;
          const double epsXmin = E[i] + E[i - JX], epsXplus = E[i] + E[i + JX];
          const double epsYmin = E[i] + E[i - 1], epsYplus = E[i] + E[i + 1];
          const double ax = x == f ? P[i - JX] : X[i - JX];
          const double ay = y == f ? P[i - 1] : X[i - 1];
          X[i] = (C * q[i] + epsXmin * ax + epsXplus * P[i + JX] + epsYmin * ay + epsYplus * P[i + 1]) /
                 (epsXmin + epsXplus + epsYmin + epsYplus);
        }
      for (int z = 0; z < M; z++) gp[z] = gp[z] - X[z];
      return;
    }
    if (s.lat.gradients == 2) {
      // LGrad2::UpdatePsi (src/LGrad2.cpp:1428-1445): Jacobi
      C = C / 2 / f / f;
      double r = off;
      for (int x = f; x < MX + f; x++) {
        r++;
        for (int y = f; y < MY + f; y++) {
          const int i = x * JX + y;
          const double Li = s.lat.L[i];
          const double epsXmin = 2 * kPIE * (r - 1) * (E[i] + E[i - JX]) / Li * f * f;
          const double epsXplus = 2 * kPIE * r * (E[i] + E[i + JX]) / Li * f * f;
          const double epsYmin = E[i] + E[i - 1], epsYplus = E[i] + E[i + 1];
          X[i] = (C * q[i] + epsXmin * P[i - JX] + epsXplus * P[i + JX] + epsYmin * P[i - 1] + epsYplus * P[i + 1]) /
                 (epsXmin + epsXplus + epsYmin + epsYplus);
        }
      }
      for (int z = 0; z < M; z++) gp[z] = gp[z] - X[z];
      return;
    }
    if (geo == GEOM_PLANAR) {
      C = C * 2.0 / f / f;
      double epsXplus = E[f - 1] + E[f];
      for (int x = f; x < MX + f; x++) {
        double epsXmin = epsXplus;
        epsXplus = E[x] + E[x + 1];
        double a = x == f ? P[f - 1] : X[x - 1];
        X[x] = (epsXmin * a + C * q[x] + epsXplus * P[x + 1]) / (epsXmin + epsXplus);
      }
    } else if (geo == GEOM_CYLINDRICAL) {
      C = C / kPIE;
      double r = off;
      double epsXplus = r * (E[f - 1] + E[f]);
      for (int x = f; x < MX + f; x++) {
        r++;
        double epsXmin = epsXplus;
        epsXplus = r * (E[x] + E[x + 1]);
        double a = x == f ? P[f - 1] : X[x - 1];
        X[x] = (epsXmin * a + C * q[x] + epsXplus * P[x + 1]) / (epsXmin + epsXplus);
      }
    } else {
      C = C / (2.0 * kPIE) * f;
      double r = off;
      double epsXplus = r * r * (E[f - 1] + E[f]);
      for (int x = f; x < MX + f; x++) {
        double epsXmin = epsXplus;
        r++;
        epsXplus = r * r * (E[x] + E[x + 1]);
        X[x] = (epsXmin * P[x - 1] + C * q[x] * s.lat.L[x] + epsXplus * P[x + 1]) / (epsXmin + epsXplus);
      }
    }
    // YisAminB over all M: the ghosts of X are never written (zero); the
    // caller removes the ghosts of g afterwards
    for (int z = 0; z < M; z++) gp[z] = gp[z] - X[z];
  }

  // Classical_residual: g from x
  void residual_eval(const double *x, double *g) {
    Lat lat(s.lat);
    const int nvar = s.nvar();
    status = 0;
    Cp(g, x, nvar);
    // PutU (src/system.cpp:1987-2085): with charges, u also carries
    // -epsilon EE + valence psi
    if (s.charged) {
      Cp(psi.data(), x + nit * M, M);
      update_EE();
    }
    for (int i = 0; i < nseg; i++)
      if (s.seg[i].field_block >= 0) {
        Cp(u[i].data(), x + s.seg[i].field_block * M, M);
        if (s.charged) {
          YplusisCtimesX(u[i].data(), EE.data(), -1.0 * s.seg[i].epsilon, M);
          if (s.seg[i].valence != 0) YplusisCtimesX(u[i].data(), psi.data(), s.seg[i].valence, M);
        }
      }
    // Segment::PrepareForCalculations (src/segment.cpp:835-879)
    for (int i = 0; i < nseg; i++) {
      const SegmentSpec &sg = s.seg[i];
      seg_phibulk[i] = 0;
      if (sg.freedom == SEG_FROZEN) {
        for (int z = 0; z < M; z++) phi[i][z] = sg.mask[z];
      } else {
        Zero(phi[i].data(), M);
      }
      if (sg.freedom == SEG_TAGGED) {  // no field; the weight is the mask, without KSAM
        for (int z = 0; z < M; z++) G1[i][z] = sg.mask[z];
        continue;
      }
      lat.set_bounds(u[i].data());
      for (int z = 0; z < M; z++) G1[i][z] = std::exp(-u[i][z]);
      if (sg.freedom == SEG_PINNED)
        for (int z = 0; z < M; z++) G1[i][z] = G1[i][z] * sg.mask[z];
      for (int z = 0; z < M; z++) G1[i][z] = G1[i][z] * s.ksam[z];
    }
    for (int i = 0; i < nmol; i++) Zero(mol_phitot[i].data(), M);
    // ComputePhis(residual) (src/system.cpp:2324-2717)
    Zero(phitot.data(), M);
    for (int i = 0; i < nseg; i++)
      if (s.seg[i].freedom == SEG_FROZEN) Add(phitot.data(), phi[i].data(), M);
    for (int i = 0; i < nmol; i++) compute_phi_mol(i);
    double A = 0, B = 0;  // bulk charge and volume fraction of free and restricted molecules
    for (int i = 0; i < nmol; i++) {
      const MoleculeSpec &m = s.mol[i];
      double nrm = 0;
      if (m.freedom == MOL_FREE) {
        nrm = phibulk[i] / m.chainlength;
        n[i] = nrm * GN[i];
        A += phibulk[i] * s.mol_charge[i];
        B += phibulk[i];
      }
      if (m.freedom == MOL_RESTRICTED) {
        if (GN[i] > 0) {
          nrm = n[i] / GN[i];
          if (m.pinned) {
            phibulk[i] = 0;
          } else {
            phibulk[i] = m.chainlength * nrm;
            A += phibulk[i] * s.mol_charge[i];
            B += phibulk[i];
          }
        } else {
          status = -1;
          return;
        }
      }
      if (m.pinned || m.tagged) {
        if (GN[i] > 0) {
          if (n[i] == 0) n[i] = 1;
          nrm = n[i] / GN[i];
        } else {
          nrm = 0;
        }
        phibulk[i] = 0;
      }
      norm[i] = nrm;
      for (size_t k = 0; k < m.molmon.size(); k++) {
        double *ph = mol_phi[i].data() + k * M;
        Div(ph, G1[m.molmon[k]].data(), M);
        if (nrm > 0) Norm(ph, nrm, M);
      }
    }
    // the neutralizer's bulk fraction makes the bulk neutral
    // (System::ComputePhis src/system.cpp:2487-2516)
    const int nz = s.charged ? s.neutralizer : -1;
    if (nz > -1) {
      const int sv = s.solvent;
      phibulk[nz] = ((B - 1.0) * s.mol_charge[sv] - A) / (s.mol_charge[nz] - s.mol_charge[sv]);
      B += phibulk[nz];
      const double nrm = phibulk[nz] / s.mol[nz].chainlength;
      n[nz] = nrm * GN[nz];
      theta[nz] = n[nz] * s.mol[nz].chainlength;
      norm[nz] = nrm;
    }
    if (s.solvent > -1) {
      const int sv = s.solvent;
      const MoleculeSpec &m = s.mol[sv];
      phibulk[sv] = 1.0 - B;
      if (phibulk[sv] < 0) {
        status = -4;
        return;
      }
      double nrm = phibulk[sv] / m.chainlength;
      n[sv] = nrm * GN[sv];
      theta[sv] = n[sv] * m.chainlength;
      norm[sv] = nrm;
      for (size_t k = 0; k < m.molmon.size(); k++)
        if (nrm > 0) Norm(mol_phi[sv].data() + k * M, nrm, M);
    }
    if (nz > -1)
      for (size_t k = 0; k < s.mol[nz].molmon.size(); k++)
        if (norm[nz] > 0) Norm(mol_phi[nz].data() + k * M, norm[nz], M);
    for (int i = 0; i < nmol; i++) {
      const MoleculeSpec &m = s.mol[i];
      for (size_t k = 0; k < m.molmon.size(); k++) {
        const double *pm = mol_phi[i].data() + k * M;
        Add(phi[m.molmon[k]].data(), pm, M);
        Add(phitot.data(), pm, M);
        Add(mol_phitot[i].data(), pm, M);
        seg_phibulk[m.molmon[k]] += fraction(i, m.molmon[k]) * phibulk[i];
      }
      if (m.freedom != MOL_RESTRICTED && m.freedom != MOL_SOLVENT) theta[i] = n[i] * m.chainlength;
    }
    // a tagged segment's density is its mask (System::ComputePhis
    // src/system.cpp:2596-2601); phitot keeps the molecule's contribution
    for (int i = 0; i < nseg; i++)
      if (s.seg[i].freedom == SEG_TAGGED)
        for (int z = 0; z < M; z++) phi[i][z] = s.seg[i].mask[z];
    for (int i = 0; i < nseg; i++) lat.set_bounds(phi[i].data());
    for (int i = 0; i < nseg; i++) lat.side(phi_side[i].data(), phi[i].data());
    // residual (src/system.cpp:2099-2157)
    Zero(alpha.data(), M);
    for (int i = 0; i < nit; i++) {
      const SegmentSpec &si = s.seg[s.itmon[i]];
      double *gi = g + i * M;
      for (int k = 0; k < nseg; k++) {
        double chi = si.chi[k];
        if (chi != 0) {
          const double *ps = phi_side[k].data();
          const double pb = seg_phibulk[k];
          for (int z = 0; z < M; z++)
            if (phitot[z] > 0) gi[z] = gi[z] - chi * (ps[z] / phitot[z] - pb);
        }
      }
    }
    for (int i = 0; i < nit; i++) Add(alpha.data(), g + i * M, M);
    Norm(alpha.data(), 1.0 / nit, M);
    for (int i = 0; i < nit; i++) {
      double *gi = g + i * M;
      for (int z = 0; z < M; z++)
        if (phitot[z] > 0)
          gi[z] = gi[z] - alpha[z] + 1 / phitot[z] - 1.0;
        else
          gi[z] = 0;
      lat.remove_bounds(gi);
      for (int z = 0; z < M; z++) gi[z] = gi[z] * s.ksam[z];
    }
    // the potential block (src/system.cpp:2161-2169)
    if (s.charged) {
      double *gp = g + nit * M;
      Cp(gp, x + nit * M, M);
      do_electrostatics(gp, x + nit * M);
      set_M_bounds(psi.data());
      psi[0] = psi[1];
      if (s.fixed_psi0) {
        // psiMask: the masks of the frozen segments with a fixed potential
        // (System::PrepareForCalculations src/system.cpp:379-388)
        std::vector<int> pmask(M, 0);
        for (int i = 0; i < nseg; i++)
          if (s.seg[i].freedom == SEG_FROZEN && s.seg[i].fixed_psi0)
            for (int z = 0; z < M; z++) pmask[z] += s.seg[i].mask[z];
        update_psi_fixed(gp, pmask);
      } else {
        update_psi(gp);
      }
      lat.remove_bounds(gp);
    }
  }

  double residual_norm(const double *g) const {
    const int nvar = s.nvar();
    if (s.solver.max_g) {
      double mx = 0;
      for (int i = 0; i < nvar; i++) mx = std::max(mx, std::fabs(g[i]));
      return mx;
    }
    return std::sqrt(Dot(g, g, nvar));
  }

  // GetFreeEnergy, GetGrandPotential, CreateMu at the state of the last residual
  void thermodynamics() {
    Lat lat(s.lat);
    vec TEMP(M, 0.0);
    for (int i = 0; i < nmol; i++) lat.remove_bounds(mol_phitot[i].data());
    for (int i = 0; i < nseg; i++) lat.remove_bounds(phi_side[i].data());
    for (int i = 0; i < nmol; i++) {
      vec t(mol_phitot[i]);
      theta_out[i] = lat.weighted_sum(t.data());
    }
    double *Fd = F_dens.data();
    Zero(Fd, M);
    for (int i = 0; i < nmol; i++) {
      // the tagged segment does not count (src/system.cpp:2861-2863)
      const int N = s.mol[i].chainlength - (s.mol[i].tagged ? 1 : 0);
      double constant = std::log(N * n[i] / GN[i]) / N;
      Cp(TEMP.data(), mol_phitot[i].data(), M);
      Norm(TEMP.data(), constant, M);
      Add(Fd, TEMP.data(), M);
    }
    // SysMonList: segments used by molecules, in order, tagged ones left
    // out (src/system.cpp:455-459)
    std::vector<int> sysmon;
    for (auto &m : s.mol)
      for (int sg : m.molmon)
        if (s.seg[sg].freedom != SEG_TAGGED && std::find(sysmon.begin(), sysmon.end(), sg) == sysmon.end())
          sysmon.push_back(sg);
    for (int j : sysmon) {
      for (int z = 0; z < M; z++) TEMP[z] = G1[j][z] > 0 ? -std::log(G1[j][z]) : 0;
      Times(TEMP.data(), phi[j].data(), TEMP.data(), M);
      Norm(TEMP.data(), -1, M);
      Add(Fd, TEMP.data(), M);
    }
    for (int j = 0; j < nseg; j++) {
      for (int k = 0; k < nseg; k++) {
        const bool fixk = s.seg[k].freedom == SEG_FROZEN || s.seg[k].freedom == SEG_TAGGED;
        double chi = fixk ? s.seg[j].chi[k] : s.seg[j].chi[k] / 2;
        if (!(s.seg[j].freedom == SEG_FROZEN || s.seg[j].freedom == SEG_TAGGED || chi == 0)) {
          Times(TEMP.data(), phi[j].data(), phi_side[k].data(), M);
          Norm(TEMP.data(), chi, M);
          Add(Fd, TEMP.data(), M);
        }
      }
    }
    for (int i = 0; i < nmol; i++) {
      const MoleculeSpec &m = s.mol[i];
      double constant = 0;
      for (int a : m.molmon)
        for (int b : m.molmon) {
          double fA = fraction(i, a), fB = fraction(i, b);
          if (m.tagged) {  // src/system.cpp:2994-3007
            const int N = m.chainlength;
            if (N > 1) {
              fA = fA * N / (N - 1);
              fB = fB * N / (N - 1);
            } else {
              fA = 0;
              fB = 0;
            }
          }
          double chi = s.seg[a].chi[b] / 2;
          constant -= fA * fB * chi;
        }
      Cp(TEMP.data(), mol_phitot[i].data(), M);
      Norm(TEMP.data(), constant, M);
      Add(Fd, TEMP.data(), M);
    }
    for (int z = 0; z < M; z++) Fd[z] = Fd[z] * s.ksam[z];
    if (s.charged)  // src/system.cpp:3019-3027: q psi / 2, frozen sites included
      for (int z = 0; z < M; z++) Fd[z] = Fd[z] + 0.5 * (q[z] * psi[z]);
    F = lat.weighted_sum(Fd);

    double *GPd = GP_dens.data();
    Zero(GPd, M);
    for (int i = 0; i < nmol; i++) {
      // tagged: N - 1 and no bulk (src/system.cpp:3053-3057)
      const int N = s.mol[i].chainlength - (s.mol[i].tagged ? 1 : 0);
      Cp(TEMP.data(), mol_phitot[i].data(), M);
      for (int z = 0; z < M; z++) TEMP[z] = TEMP[z] + (-phibulk[i]);
      Norm(TEMP.data(), 1.0 / N, M);
      Add(GPd, TEMP.data(), M);
    }
    Add(GPd, alpha.data(), M);
    for (int i = 0; i < nseg; i++) lat.remove_bounds(phi_side[i].data());
    for (int j = 0; j < nseg; j++) {
      if (s.seg[j].freedom == SEG_FROZEN || s.seg[j].freedom == SEG_TAGGED) continue;
      double pbA = seg_phibulk[j];
      for (int k = 0; k < nseg; k++) {
        if (s.seg[k].freedom == SEG_FROZEN || s.seg[k].freedom == SEG_TAGGED) continue;
        double chi = s.seg[j].chi[k] / 2;
        double pbB = seg_phibulk[k];
        if (chi != 0) {
          Times(TEMP.data(), phi[j].data(), phi_side[k].data(), M);
          for (int z = 0; z < M; z++) TEMP[z] = TEMP[z] + (-pbA * pbB);
          Norm(TEMP.data(), chi, M);
          Add(GPd, TEMP.data(), M);
        }
      }
    }
    // u_ext is zero for every segment on this path (no external potentials)
    Norm(GPd, -1.0, M);
    if (s.charged) {
      // src/system.cpp:3223-3238: eps EE - q psi / 2 on accessible sites,
      // q psi / 2 on the others
      for (int z = 0; z < M; z++) {
        double t = EE[z] * eps[z];
        t = t * -2.0;
        t = t + q[z] * psi[z];
        t = t * (-1.0 / 2.0);
        GPd[z] = (GPd[z] + t) * s.ksam[z];
        GPd[z] = GPd[z] + (q[z] - q[z] * s.ksam[z]) * psi[z] * 0.5;
      }
    } else {
      for (int z = 0; z < M; z++) GPd[z] = GPd[z] * s.ksam[z];
    }
    GP = lat.weighted_sum(GPd);

    for (int i = 0; i < nmol; i++) {
      // a tagged molecule counts N - 1 segments; its tagged segment has no
      // fraction and the others are scaled to N / (N - 1)
      // (src/system.cpp:3262-3340)
      const bool tg = s.mol[i].tagged;
      const double NA = s.mol[i].chainlength - (tg ? 1 : 0);
      double mu = std::log(NA * n[i] / GN[i]) + 1;
      double constant = 0;
      for (int k = 0; k < nmol; k++) {
        const double NB = s.mol[k].chainlength - (s.mol[k].tagged ? 1 : 0);
        constant += phibulk[k] / NB;
      }
      mu = mu - NA * constant;
      for (int j = 0; j < nseg; j++) {
        double pbA = seg_phibulk[j];
        double FA = s.seg[j].freedom == SEG_TAGGED ? 0 : fraction(i, j);
        if (tg) FA *= (NA + 1) / (NA);
        for (int k = 0; k < nseg; k++) {
          double pbB = seg_phibulk[k];
          double FB = s.seg[k].freedom == SEG_TAGGED ? 0 : fraction(i, k);
          if (tg) FB *= (NA + 1) / (NA);
          double chi = s.seg[j].chi[k] / 2;
          mu = mu - NA * chi * (pbA - FA) * (pbB - FB);
        }
      }
      Mu[i] = mu;
    }
  }

  // SFNewton::iterate_DIIS (src/sfnewton.cpp:1037-1130)
  int solve_diis(double *x, int &iterations, double &res) {
    const int nvar = s.nvar();
    const int m = s.solver.m;
    const double dmax = s.solver.deltamax, tol = s.solver.tolerance;
    const int itlim = s.solver.iterationlimit;
    vec Aij(m * m, 0.0), Apij(m * m, 0.0), Ci(m, 0.0);
    vec xR((size_t)m * nvar, 0.0), x_x0((size_t)m * nvar, 0.0), x0(nvar, 0.0), g(nvar, 0.0);
    iterations = 0;
    int k_diis = 0, k = 0;
    Cp(x0.data(), x, nvar);
    residual_eval(x, g.data());
    if (status) return status;
    YplusisCtimesX(x, g.data(), -dmax, nvar);
    for (int i = 0; i < nvar; i++) x_x0[i] = x[i] - x0[i];
    Cp(xR.data(), x, nvar);
    res = residual_norm(g.data());
    while (res > tol && iterations < itlim) {
      iterations++;
      Cp(x0.data(), x, nvar);
      residual_eval(x, g.data());
      if (status) return status;
      k = iterations % m;
      k_diis++;
      YplusisCtimesX(x, g.data(), -dmax, nvar);
      Cp(xR.data() + (size_t)k * nvar, x, nvar);
      for (int i = 0; i < nvar; i++) x_x0[(size_t)k * nvar + i] = x[i] - x0[i];
      // DIIS (src/sfnewton.cpp:837-880)
      int kd = k_diis;
      if (kd > m) {
        kd = m;
        for (int i = 1; i < m; i++)
          for (int j = 1; j < m; j++) Aij[m * (i - 1) + j - 1] = Aij[m * i + j];
      }
      for (int i = 0; i < kd; i++) {
        int posi = k - kd + 1 + i;
        if (posi < 0) posi += m;
        double D = Dot(x_x0.data() + (size_t)posi * nvar, x_x0.data() + (size_t)k * nvar, nvar);
        Aij[i + m * (kd - 1)] = Aij[kd - 1 + m * i] = D + 1e-9;
        for (int j = 0; j < kd; j++) Apij[j + kd * i] = Aij[j + m * i];
      }
      vec Apc(Apij.begin(), Apij.begin() + kd * kd), C;
      if (!solve_ones(Apc, kd, C)) {
        status = -2;
        return status;
      }
      double normC = 0;
      for (int i = 0; i < kd; i++) normC += C[i];
      for (int i = 0; i < kd; i++) C[i] = C[i] / normC;
      Zero(x, nvar);
      int posi = k - kd + 1;
      if (posi < 0) posi += m;
      YplusisCtimesX(x, xR.data() + (size_t)posi * nvar, C[0], nvar);
      for (int i = 1; i < kd; i++) {
        posi = k - kd + 1 + i;
        if (posi < 0) posi += m;
        YplusisCtimesX(x, xR.data() + (size_t)posi * nvar, C[i], nvar);
      }
      res = residual_norm(g.data());
    }
    return res <= tol ? 0 : 1;
  }
};

}  // namespace

// ----------------------------------------------------------------- C ABI
extern "C" {

struct oracle_handle {
  Oracle *o;
};

void *oracle_create(const char *text, int start, const char *guess_text, char *err, int errlen) {
  SystemSpec s;
  std::string e;
  if (!parse_input(text, start, s, e)) {
    if (err && errlen > 0) snprintf(err, errlen, "%s", e.c_str());
    return nullptr;
  }
  (void)guess_text;
  return new Oracle(s);
}

void oracle_destroy(void *h) { delete (Oracle *)h; }

// JSON description of the parsed system; returns the length needed.
int oracle_describe(void *h, char *buf, int len) {
  std::string d = describe_json(((Oracle *)h)->s);
  if (buf && len > 0) snprintf(buf, len, "%s", d.c_str());
  return (int)d.size() + 1;
}

// sizes: [M, nvar, nseg, nmol, MX, MY, MZ, gradients, nit]
void oracle_sizes(void *h, int *out) {
  Oracle *o = (Oracle *)h;
  const LatticeSpec &l = o->s.lat;
  int v[9] = {o->M, o->s.nvar(), o->nseg, o->nmol, l.MX, l.MY, l.MZ, l.gradients, o->nit};
  std::memcpy(out, v, sizeof(v));
}

int oracle_read_guess(void *h, const char *guess_text, double *x, char *err, int errlen) {
  Oracle *o = (Oracle *)h;
  std::vector<double> xv;
  std::string e;
  if (!read_guess_text(guess_text, o->s, xv, e)) {
    if (err && errlen > 0) snprintf(err, errlen, "%s", e.c_str());
    return -1;
  }
  std::copy(xv.begin(), xv.end(), x);
  return 0;
}

// Electrostatic profiles of the last residual: psi, q, eps, EE (4 rows of M).
void oracle_electro(void *h, double *out) {
  Oracle *o = (Oracle *)h;
  const int M = o->M;
  std::copy(o->psi.begin(), o->psi.end(), out);
  std::copy(o->q.begin(), o->q.end(), out + M);
  std::copy(o->eps.begin(), o->eps.end(), out + 2 * M);
  std::copy(o->EE.begin(), o->EE.end(), out + 3 * M);
}

// Residual at x; fills g and keeps the state for the getters below.
int oracle_residual(void *h, const double *x, double *g, double *res) {
  Oracle *o = (Oracle *)h;
  o->residual_eval(x, g);
  if (o->status) return o->status;
  *res = o->residual_norm(g);
  o->thermodynamics();
  return 0;
}

int oracle_solve(void *h, double *x, int *iterations, double *res) {
  Oracle *o = (Oracle *)h;
  int it = 0;
  double r = 0;
  int st = o->solve_diis(x, it, r);
  *iterations = it;
  *res = r;
  if (st < 0) return st;
  o->thermodynamics();
  return st;
}

// scalars: [F, GP] then per molecule [GN, n, theta, phibulk, mu, norm]
void oracle_scalars(void *h, double *out) {
  Oracle *o = (Oracle *)h;
  out[0] = o->F;
  out[1] = o->GP;
  for (int i = 0; i < o->nmol; i++) {
    double *p = out + 2 + 6 * i;
    p[0] = o->GN[i];
    p[1] = o->n[i];
    p[2] = o->theta_out[i];
    p[3] = o->phibulk[i];
    p[4] = o->Mu[i];
    p[5] = o->norm[i];
  }
}

// per-segment profiles (nseg x M): phi, and G1
void oracle_seg_phi(void *h, double *out) {
  Oracle *o = (Oracle *)h;
  for (int i = 0; i < o->nseg; i++) std::copy(o->phi[i].begin(), o->phi[i].end(), out + (size_t)i * o->M);
}

void oracle_seg_g1(void *h, double *out) {
  Oracle *o = (Oracle *)h;
  for (int i = 0; i < o->nseg; i++) std::copy(o->G1[i].begin(), o->G1[i].end(), out + (size_t)i * o->M);
}

// per-molecule total density (nmol x M), ghosts removed
void oracle_mol_phi(void *h, double *out) {
  Oracle *o = (Oracle *)h;
  for (int i = 0; i < o->nmol; i++)
    std::copy(o->mol_phitot[i].begin(), o->mol_phitot[i].end(), out + (size_t)i * o->M);
}

// Plain propagation of one chain of length N in the uniform field
// G1 (M values): forward end-point distributions out[N*M]
// (used by closed-form tests of the stencils).
int oracle_propagate_chain(void *h, const double *G1, int N, double *out) {
  Oracle *o = (Oracle *)h;
  Lat lat(o->s.lat);
  const int M = o->M;
  Cp(out, G1, M);
  for (int k = 1; k < N; k++) lat.propagate(out + (size_t)k * M, out + (size_t)(k - 1) * M, G1);
  return 0;
}

// One application of the lattice operator to X (M values, bounded in place
// like the reference does): side=0 the propagator stencil with G1 = 1,
// side=1 the site average used for the interaction term.
int oracle_apply(void *h, double *X, double *out, int side) {
  Oracle *o = (Oracle *)h;
  Lat lat(o->s.lat);
  if (side) {
    lat.set_bounds(X);
    lat.side(out, X);
  } else {
    std::vector<double> one(o->M, 1.0);
    lat.propagate(out, X, one.data());
  }
  return 0;
}

// Markov-2 forward step (1 gradient): in/out hold three components (3M).
int oracle_apply_markov2(void *h, double *X, double *out, int mol, int forward) {
  Oracle *o = (Oracle *)h;
  Lat lat(o->s.lat);
  std::vector<double> one(o->M, 1.0);
  const double *P = o->s.mol[mol].P.data();
  if (forward)
    lat.propagateF(out, X, one.data(), P);
  else
    lat.propagateB(out, X, one.data(), P);
  return 0;
}

// A c = 1 by the SVD path (for tests of the DIIS least-squares step)
int oracle_solve_ones(const double *A, int N, double *c) {
  vec Av(A, A + N * N), C;
  if (!solve_ones(Av, N, C)) return -1;
  std::copy(C.begin(), C.end(), c);
  return 0;
}
}
