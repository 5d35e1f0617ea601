// Propagator schedule of a branched molecule.
//
// The reference computes a branched chain (square brackets in the
// composition) by a recursion over its generations:
// mol_branched::ForwardBra / BackwardBra (src/mol_branched.cpp:166-300) on
// top of Molecule::propagate_forward / propagate_backward
// (src/molecule.cpp:2300-2446), driven by mol_branched::ComputePhi
// (:303-420).  The recursion depends only on the composition, so it is
// unrolled here, once, on the host, into a flat list of row operations that
// the device executes in order (scf_grid.hip).  The molecule is the one the
// parser leaves (turned around as the reference turns it: blocks from the
// main chain's far end, first_s / last_s the first and last segment of each
// generation).  The schedule follows the reference step by step, including
// one corner of it: a generation whose last block is a branch point returns
// the end row of the segment in front of its side groups, so a composition
// that starts with a one-segment branch point sums GN over that row.  The
// Markov-1 Initiate and Terminate are plain copies (LGrad1::Initiate /
// Terminate src/LGrad1.cpp:943-969), and every scratch array the reference
// allocates per call (GS, GX) is a row of its own per recursion depth, so
// nested calls keep their caller's rows.
#include <algorithm>
#include <array>
#include <functional>
#include <string>
#include <vector>

#include "scf_spec.h"

namespace nscf {

namespace {

struct PlanBuilder {
  const MoleculeSpec &m;
  BranchPlan &plan;
  std::string &err;
  int cl = 0;                     // chain length: rows 0 .. cl-1 are the forward rows
  int gb0 = 0;                    // backward rows gb0, gb0 + 1 (Gg_b, slots s % 2)
  std::vector<std::vector<int>> gs;  // scratch rows per recursion depth
  int next_row = 0;
  bool ok = true;

  int scratch(int depth, int k) {
    while ((int)gs.size() <= depth) gs.emplace_back();
    while ((int)gs[depth].size() <= k) gs[depth].push_back(next_row++);
    return gs[depth][k];
  }
  void emit(std::vector<BranchOp> &ops, int kind, int a, int b, int c, int seg) { ops.push_back({kind, a, b, c, seg}); }
  int seg_of(int block) const { return m.blocks[block].seg; }
  int mm_of(int block) const { return m.block_molmon[block]; }

  // Molecule::propagate_forward without save_memory: rows s, s+1, ... of
  // one block; the first row of a generation is its segment weight.
  int forward_block(int block, int gen, int &s) {
    for (int k = 0; k < m.blocks[block].n; k++) {
      if (s > m.first_s[gen])
        emit(plan.forward, BR_PROP, s, s - 1, 0, seg_of(block));
      else
        emit(plan.forward, BR_INIT, m.first_s[gen], 0, 0, seg_of(block));
      s++;
    }
    return s - 1;
  }

  // mol_branched::ForwardBra: returns the row the reference returns (-1 for
  // its null pointer).
  int forward(int gen, int &s, int depth) {
    const int b0 = m.first_b[gen], bN = m.last_b[gen];
    int glast = -1;
    for (int k = b0; k <= bN && ok; ++k) {
      if (b0 < k && k < bN && m.gnr[k] != gen) {
        // a junction: side groups start at block k
        const int g0 = scratch(depth, 0), g1 = scratch(depth, 1), g2 = scratch(depth, 2);
        if (glast < 0) {
          err = "branched composition: a side group with no segment in front of it";
          ok = false;
          return -1;
        }
        emit(plan.forward, BR_COPY, g0, glast, 0, 0);
        std::vector<int> ends;
        while (k <= bN && m.gnr[k] != gen && ok) {
          const int g = m.gnr[k];
          ends.push_back(forward(g, s, depth + 1));
          k += m.last_b[g] - m.first_b[g] + 1;
        }
        if (!ok) return -1;
        if (k > bN) {
          err = "branched composition: a side group at the end of a generation";
          ok = false;
          return -1;
        }
        // the segment after the group: backbone step times every side chain
        // end stepped once without a segment weight
        emit(plan.forward, BR_PROP, g2, g0, 0, seg_of(k));
        for (int e : ends) {
          if (e < 0) {
            err = "branched composition: a side group without an end row";
            ok = false;
            return -1;
          }
          emit(plan.forward, BR_COPY, g0, e, 0, 0);
          emit(plan.forward, BR_PROP, g1, g0, 0, -1);
          emit(plan.forward, BR_TIMES, g2, g2, g1, 0);
        }
        emit(plan.forward, BR_COPY, s, g2, 0, 0);
        s++;
      } else {
        glast = forward_block(k, gen, s);
      }
    }
    return glast;
  }

  // Molecule::propagate_backward without save_memory.
  void backward_block(int block, int &s) {
    for (int k = 0; k < m.blocks[block].n; k++) {
      const int cur = gb0 + ((s % 2) + 2) % 2, prev = gb0 + (((s + 1) % 2) + 2) % 2;
      if (s < cl - 1)
        emit(plan.backward, BR_PROP, cur, prev, 0, seg_of(block));
      else
        emit(plan.backward, BR_INIT, cur, 0, 0, seg_of(block));
      if (s < 0 || s >= cl) {
        err = "branched composition: the backward pass leaves the chain";
        ok = false;
        return;
      }
      emit(plan.backward, BR_ADDPHI, s, cur, mm_of(block), 0);
      s--;
    }
  }

  // mol_branched::BackwardBra
  void backward(int gen, int &s, int depth) {
    const int b0 = m.first_b[gen], bN = m.last_b[gen];
    for (int k = bN; k >= b0 && ok; k--) {
      if (m.gnr[k] != gen) {
        std::vector<int> br, rows;
        int ss = 0;
        while (k >= 0 && m.gnr[k] != gen) {
          const int g = m.gnr[k];
          br.push_back(g);
          rows.push_back(m.last_s[g]);
          ss = m.first_s[g];
          k -= m.last_b[g] - m.first_b[g] + 1;
        }
        br.push_back(gen);
        ss--;
        rows.push_back(ss);
        const int L = (int)br.size();
        for (int r : rows)
          if (r < 0 || r >= cl) {
            err = "branched composition: a junction reads a row outside the chain";
            ok = false;
            return;
          }
        const int g0 = scratch(depth, 0), g1 = scratch(depth, 1), g2 = scratch(depth, 2), g3 = scratch(depth, 3);
        emit(plan.backward, BR_COPY, g3, gb0 + (((s + 1) % 2) + 2) % 2, 0, 0);
        for (int i = 0; i < L && ok; i++) {
          emit(plan.backward, BR_COPY, g2, g3, 0, 0);
          for (int j = 0; j < L; j++) {
            if (j == i) continue;
            emit(plan.backward, BR_COPY, g0, rows[j], 0, 0);
            emit(plan.backward, BR_PROP, g1, g0, 0, -1);
            emit(plan.backward, BR_TIMES, g2, g2, g1, 0);
          }
          emit(plan.backward, BR_COPY, gb0, g2, 0, 0);
          emit(plan.backward, BR_COPY, gb0 + 1, g2, 0, 0);
          if (i < L - 1) backward(br[i], s, depth + 1);
        }
        k++;
      } else {
        backward_block(k, s);
      }
    }
  }
};

}  // namespace

// mol_dend::BackAndForth (src/mol_dendrimer.cpp:122-188): one path from the
// core through one arm of every generation, rows 0 .. slast (row 0 the core's
// branch point).  Forward from the outer end: each generation's spacer, then
// its branch point as its weight times the spacer's first row stepped once,
// to the power of the functionality.  Backward from the core: the core's
// weight times all but one of its arms, then along the path, every
// composition weighted by the block's degeneracy; at each further branch
// point the backward row takes the other arms of that generation.
bool build_dendrimer_plan(const MoleculeSpec &m, BranchPlan &plan, std::string &err) {
  const int G = (int)m.n_arm.size();
  if (G == 0 || (int)m.first_b.size() != G || (int)m.last_s.size() != G) {
    err = "dendrimer without generations";
    return false;
  }
  const int slast = m.last_s[G - 1];
  const int gb0 = slast + 1, g0 = slast + 3, g1 = slast + 4, g2 = slast + 5;
  std::vector<BranchOp> &fw = plan.forward, &bw = plan.backward;
  auto op = [](std::vector<BranchOp> &v, int kind, int a, int b, int c, int seg, double f = 1.0) {
    v.push_back({kind, a, b, c, seg, f});
  };
  auto seg_of = [&](int b) { return m.blocks[b].seg; };
  auto mm_of = [&](int b) { return m.block_molmon[b]; };
  auto gb = [&](int s) { return gb0 + (((s % 2) + 2) % 2); };
  int s = slast;
  for (int g = G - 1; g >= 0; g--) {
    op(fw, BR_ONE, g2, 0, 0, 0);
    const int b0 = m.first_b[g], bN = m.last_b[g];
    for (int b = bN; b >= b0; b--)
      for (int k = 0; k < m.blocks[b].n; k++) {
        if (s < slast)
          op(fw, BR_PROP, s, s + 1, 0, seg_of(b));
        else
          op(fw, BR_INIT, slast, 0, 0, seg_of(b));
        s--;
      }
    if (s < 0 || b0 < 1) {
      err = "dendrimer path out of range";
      return false;
    }
    op(fw, BR_COPY, g0, s + 1, 0, 0);
    op(fw, BR_PROP, g1, g0, 0, -1);
    op(fw, BR_PROP, g2, g0, 0, seg_of(b0 - 1));
    for (int k = 0; k < m.n_arm[g] - 1; k++) op(fw, BR_TIMES, g2, g2, g1, 0);
    op(fw, BR_COPY, s, g2, 0, 0);
    s--;
  }
  if (s != -1) {
    err = "dendrimer path does not end at the core";
    return false;
  }
  plan.gn_row = 0;
  op(bw, BR_INIT, gb0, 0, 0, seg_of(0));
  op(bw, BR_ADDPHI, 0, gb0, mm_of(0), 0);
  s = 0;
  op(bw, BR_ONE, g2, 0, 0, 0);
  op(bw, BR_COPY, g0, 1, 0, 0);
  op(bw, BR_PROP, g1, g0, 0, -1);
  for (int k = 0; k < m.n_arm[0] - 1; k++) op(bw, BR_TIMES, g2, g2, g1, 0);
  op(bw, BR_TIMES_G1, g2, g2, 0, seg_of(0));
  op(bw, BR_COPY, gb0, g2, 0, 0);
  for (int g = 0; g < G; g++) {
    op(bw, BR_ONE, g2, 0, 0, 0);
    const int b0 = m.first_b[g], bN = m.last_b[g];
    for (int b = b0; b <= bN; b++)
      for (int k = 0; k < m.blocks[b].n; k++) {
        op(bw, BR_PROP, gb(s + 1), gb(s), 0, seg_of(b));
        s++;
        op(bw, BR_ADDPHI, s, gb(s), mm_of(b), 0, (double)m.block_deg[b]);
      }
    if (s < slast) {
      const int bp = bN + 1;  // the next generation's branch point
      op(bw, BR_ONE, g2, 0, 0, 0);
      op(bw, BR_COPY, g0, s + 2, 0, 0);
      op(bw, BR_PROP, g1, g0, 0, -1);
      for (int k = 0; k < m.n_arm[g + 1] - 1; k++) op(bw, BR_TIMES, g2, g2, g1, 0);
      op(bw, BR_PROP, gb(s + 1), gb(s), 0, seg_of(bp));
      s++;
      op(bw, BR_ADDPHI, s, gb(s), mm_of(bp), 0, (double)m.block_deg[bp]);
      op(bw, BR_COPY, g0, gb(s), 0, 0);
      op(bw, BR_TIMES, g0, g0, g2, 0);
      op(bw, BR_COPY, gb(s), g0, 0, 0);
    }
  }
  plan.rows = g2 + 1;
  return true;
}

// mol_comb::GoBackAndForth (src/mol_comb.cpp:206-400), no ring.  Rows are
// path indices as the reference uses them: the tooth is propagated once,
// from its free end, into rows first_s[1]+1 .. last_s[1]+1, and its row next
// to the backbone, stepped once without a weight, is the factor GB every
// branch point carries; the backbone runs from the tail's end to the head,
// the first branch point stored at row last_s[0]+1.  The backward pass
// walks the backbone, turns each branch point's forward row into the
// tooth's backward start (Gb Gf / GB / G1) and, last, propagates the tooth
// once per tooth from those rows.
// mol_asym_dend::Forward / Backward (src/mol_asym_dendrimer.cpp:127-229),
// several arm types per generation.  Forward from the outer generation:
// each arm type's spacer from the next generation's branch point, and the
// branch point as its weight times every arm's first row stepped once, to
// the power of that arm type's count.  Backward by recursion: at a branch
// point of generation g (reached n_repeats times), GX = Gb Gf / G1; each
// arm type starts from GX divided by its own stepped first row, and the next
// generation is walked again for every arm type of this one (compositions
// weighted by n_arm n_repeats).  GX and the stepped rows are per recursion
// depth, as the reference allocates them per call.
bool build_asym_dendrimer_plan(const MoleculeSpec &m, BranchPlan &plan, std::string &err) {
  const int G = (int)m.first_a.size();
  int maxrow = 0;
  for (int l : m.last_s) maxrow = std::max(maxrow, l);
  const int gb0 = maxrow + 1;
  const int fg0 = maxrow + 3, fg1 = maxrow + 4, fg2 = maxrow + 5;  // forward GS rows
  int next_row = maxrow + 6;
  std::vector<std::array<int, 3>> depth_rows;  // per generation: GX, GS, GS+M
  for (int g = 0; g < G; g++) {
    depth_rows.push_back({next_row, next_row + 1, next_row + 2});
    next_row += 3;
  }
  std::vector<BranchOp> &fw = plan.forward, &bw = plan.backward;
  auto op = [](std::vector<BranchOp> &v, int kind, int a, int b, int c, int seg, double f = 1.0) {
    v.push_back({kind, a, b, c, seg, f});
  };
  auto seg_of = [&](int b) { return m.blocks[b].seg; };
  auto mm_of = [&](int b) { return m.block_molmon[b]; };
  auto gb = [&](int s) { return gb0 + (((s % 2) + 2) % 2); };
  int slast = 0, s = 0;
  for (int g = G - 1; g >= 0; g--) {
    s = m.last_s[m.last_a[g]];
    op(fw, BR_ONE, fg2, 0, 0, 0);
    for (int a = m.last_a[g]; a >= m.first_a[g]; a--) {
      const int b0 = m.first_b[a], bN = m.last_b[a];
      for (int b = bN; b >= b0; b--)
        for (int k = 0; k < m.blocks[b].n; k++) {
          if (g == G - 1 && s == m.last_s[a])
            op(fw, BR_INIT, s, 0, 0, seg_of(b));
          else if (b == bN && k == 0)
            op(fw, BR_PROP, s, slast, 0, seg_of(b));
          else
            op(fw, BR_PROP, s, s + 1, 0, seg_of(b));
          s--;
        }
      op(fw, BR_COPY, fg0, s + 1, 0, 0);
      op(fw, BR_PROP, fg1, fg0, 0, -1);
      for (int k = 0; k < m.n_arm[a]; k++) op(fw, BR_TIMES, fg2, fg2, fg1, 0);
    }
    op(fw, BR_TIMES_G1, fg2, fg2, 0, seg_of(m.first_b[m.first_a[g]] - 1));
    if (s < 0) {
      err = "dendrimer path out of range";
      return false;
    }
    op(fw, BR_COPY, s, fg2, 0, 0);
    slast = s;
  }
  if (s != 0) {
    err = "dendrimer path does not end at the core";
    return false;
  }
  plan.gn_row = 0;
  // the backward recursion
  std::function<void(int, int)> backward = [&](int g, int nrep) {
    const int GX = depth_rows[g][0], GS0 = depth_rows[g][1], GS1 = depth_rows[g][2];
    const int bp = m.first_b[m.first_a[g]] - 1;
    int s = 0;
    if (g == 0) {
      op(bw, BR_INIT, gb(0), 0, 0, seg_of(bp));
    } else {
      s = m.first_s[m.first_a[g]] - 1;
      op(bw, BR_PROP, gb(s), gb(s - 1), 0, seg_of(bp));
    }
    op(bw, BR_TIMES, GX, gb(s), s, 0);
    op(bw, BR_ADDPHI, s, gb(s), mm_of(bp), 0, (double)nrep);
    op(bw, BR_DIV_G1, GX, 0, 0, seg_of(bp));
    for (int a = m.first_a[g]; a <= m.last_a[g]; a++) {
      s = m.first_s[a];
      op(bw, BR_COPY, GS0, s, 0, 0);
      op(bw, BR_PROP, GS1, GS0, 0, -1);
      op(bw, BR_COPY, gb(s - 1), GX, 0, 0);
      op(bw, BR_DIV, gb(s - 1), GS1, 0, 0);
      for (int b = m.first_b[a]; b <= m.last_b[a]; b++)
        for (int k = 0; k < m.blocks[b].n; k++) {
          op(bw, BR_PROP, gb(s), gb(s - 1), 0, seg_of(b));
          op(bw, BR_ADDPHI, s, gb(s), mm_of(b), 0, (double)m.n_arm[a] * nrep);
          s++;
        }
      op(bw, BR_COPY, gb(s), gb(s - 1), 0, 0);
      if (g < G - 1) backward(g + 1, m.n_arm[a] * nrep);
    }
  };
  backward(0, 1);
  plan.rows = next_row;
  return true;
}

bool build_comb_plan(const MoleculeSpec &m, BranchPlan &plan, std::string &err) {
  const int narm = m.n_arm.empty() ? 0 : m.n_arm[0];
  const int G = narm + 3;
  if (narm < 1 || (int)m.first_b.size() != G) {
    err = "comb without teeth";
    return false;
  }
  int maxrow = 0;
  for (int g = 0; g < G; g++) maxrow = std::max(maxrow, m.last_s[g] + 1);
  const int gb0 = maxrow + 1, GB0 = maxrow + 3, GB1 = maxrow + 4;
  std::vector<BranchOp> &fw = plan.forward, &bw = plan.backward;
  auto op = [](std::vector<BranchOp> &v, int kind, int a, int b, int c, int seg) { v.push_back({kind, a, b, c, seg}); };
  auto seg_of = [&](int b) { return m.blocks[b].seg; };
  auto mm_of = [&](int b) { return m.block_molmon[b]; };
  auto gb = [&](int s) { return gb0 + (((s % 2) + 2) % 2); };
  // the tooth, from its free end
  int s = m.last_s[1] + 1, slast = s;
  for (int b = m.last_b[1]; b >= m.first_b[1]; b--)
    for (int k = 0; k < m.blocks[b].n; k++) {
      if (s < slast)
        op(fw, BR_PROP, s, s + 1, 0, seg_of(b));
      else
        op(fw, BR_INIT, slast, 0, 0, seg_of(b));
      s--;
    }
  op(fw, BR_COPY, GB1, s + 1, 0, 0);
  op(fw, BR_PROP, GB0, GB1, 0, -1);
  const int sfirst = s + 1;
  // the tail, from the backbone's end
  s = m.last_s[narm + 2];
  slast = s;
  for (int b = m.last_b[narm + 2]; b >= m.first_b[narm + 2]; b--)
    for (int k = 0; k < m.blocks[b].n; k++) {
      if (s < slast)
        op(fw, BR_PROP, s, s + 1, 0, seg_of(b));
      else
        op(fw, BR_INIT, slast, 0, 0, seg_of(b));
      s--;
    }
  // the repeats: spacer, then the branch point with its tooth
  for (int g = narm + 1; g >= 2; g--) {
    const int b0 = m.first_b[g];
    for (int b = m.last_b[g]; b >= b0; b--)
      for (int k = 0; k < m.blocks[b].n; k++) {
        op(fw, BR_PROP, s, s + 1, 0, seg_of(b));
        s--;
      }
    const int row = g == 2 ? m.last_s[0] + 1 : s;
    op(fw, BR_PROP, row, s + 1, 0, seg_of(b0 - 1));
    op(fw, BR_TIMES, row, row, GB0, 0);
    s--;
  }
  // the head
  s = m.last_s[0];
  for (int b = m.last_b[0]; b >= m.first_b[0]; b--)
    for (int k = 0; k < m.blocks[b].n; k++) {
      op(fw, BR_PROP, s, s + 1, 0, seg_of(b));
      s--;
    }
  if (s != -1) {
    err = "comb path does not end at the head";
    return false;
  }
  plan.gn_row = 0;
  op(bw, BR_INIT, gb0, 0, 0, seg_of(0));
  op(bw, BR_ADDPHI, 0, gb0, mm_of(0), 0);
  s = -1;
  for (int b = m.first_b[0]; b <= m.last_b[0]; b++)
    for (int k = 0; k < m.blocks[b].n; k++) {
      if (s >= 0) {
        op(bw, BR_PROP, gb(s + 1), gb(s), 0, seg_of(b));
        op(bw, BR_ADDPHI, s + 1, gb(s + 1), mm_of(b), 0);
      }
      s++;
    }
  for (int g = 2; g <= narm + 1; g++) {
    const int bp = m.first_b[g] - 1;
    op(bw, BR_PROP, gb(s + 1), gb(s), 0, seg_of(bp));
    op(bw, BR_ADDPHI, gb(s + 1), s + 1, mm_of(bp), 0);
    op(bw, BR_TIMES, s + 1, gb(s + 1), s + 1, 0);
    op(bw, BR_DIV, s + 1, GB0, 0, 0);
    op(bw, BR_DIV_G1, s + 1, 0, 0, seg_of(bp));
    op(bw, BR_TIMES, gb(m.first_s[g] - 1), gb(s + 1), GB0, 0);
    s = m.first_s[g] - 1;
    for (int b = m.first_b[g]; b <= m.last_b[g]; b++)
      for (int k = 0; k < m.blocks[b].n; k++) {
        op(bw, BR_PROP, gb(s + 1), gb(s), 0, seg_of(b));
        op(bw, BR_ADDPHI, s + 1, gb(s + 1), mm_of(b), 0);
        s++;
      }
  }
  for (int b = m.first_b[narm + 2]; b <= m.last_b[narm + 2]; b++)
    for (int k = 0; k < m.blocks[b].n; k++) {
      op(bw, BR_PROP, gb(s + 1), gb(s), 0, seg_of(b));
      op(bw, BR_ADDPHI, s + 1, gb(s + 1), mm_of(b), 0);
      s++;
    }
  // the teeth
  for (int g = 2; g <= narm + 1; g++) {
    s = sfirst - 1;
    op(bw, BR_COPY, GB0, g == 2 ? m.last_s[0] + 1 : m.first_s[g] - 1, 0, 0);
    const int b0 = m.first_b[1];
    for (int b = b0; b <= m.last_b[1]; b++)
      for (int k = 0; k < m.blocks[b].n; k++) {
        if (k == 0 && b == b0) {
          op(bw, BR_PROP, GB1, GB0, 0, seg_of(b));
          op(bw, BR_COPY, gb(s + 1), GB1, 0, 0);
        } else {
          op(bw, BR_PROP, gb(s + 1), gb(s), 0, seg_of(b));
        }
        op(bw, BR_ADDPHI, s + 1, gb(s + 1), mm_of(b), 0);
        s++;
      }
  }
  plan.rows = GB1 + 1;
  for (const auto &v : {plan.forward, plan.backward})
    for (const BranchOp &o : v)
      if (o.a < 0 || o.a >= plan.rows || o.b < 0 || o.b >= plan.rows || (o.kind == BR_TIMES && o.c >= plan.rows)) {
        err = "comb schedule reads a row outside the plan";
        return false;
      }
  return true;
}

bool build_branch_plan(const MoleculeSpec &m, BranchPlan &plan, std::string &err) {
  plan = BranchPlan();
  if (m.dendrimer) return m.asymmetric ? build_asym_dendrimer_plan(m, plan, err) : build_dendrimer_plan(m, plan, err);
  if (m.comb) return build_comb_plan(m, plan, err);
  if (!m.branched) {
    err = "not a branched molecule";
    return false;
  }
  PlanBuilder b{m, plan, err};
  b.cl = m.chainlength;
  b.gb0 = m.chainlength;
  b.next_row = m.chainlength + 2;
  int s = 0;
  const int g = b.forward(0, s, 0);
  if (!b.ok) return false;
  // every contour row must have been written before the backward pass reads
  // it: the block after each side group is one segment long
  if (s != m.chainlength || g < 0) {
    err = "branched composition: the segment after each side group must be a block of its own, (X)1 "
          "(the reference's propagation writes one row there)";
    return false;
  }
  plan.gn_row = g;
  s--;
  b.backward(0, s, 0);
  if (!b.ok) return false;
  plan.rows = b.next_row;
  return true;
}

}  // namespace nscf
