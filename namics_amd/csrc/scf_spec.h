// System description for the lattice self-consistent-field (SF-SCF) hot path.
//
// One SystemSpec is everything the residual and solver need about one SCF
// problem: the lattice (geometry, boundary conditions, lambda weights), the
// segment types (freedoms, masks, chi matrix), the chain molecules (block
// sequences, normalisation mode, Markov order) and the solver settings.  It
// is built from the reference's `.in` text format by parse_input() so the
// same input files drive the CPU checker and the GPU path.
//
// Reference counterparts: Lattice::CheckInput / AllocateMemory / ComputeLambdas
// (src/lattice.cpp:267-601, src/LGrad1.cpp:16-177, src/LGrad2.cpp:13-100,
// src/LG1Planar.cpp:11-26), Segment::ParseFreedoms (src/segment.cpp:156-700),
// Molecule::CheckInput (src/molecule.cpp:356-677), System::CheckChi_values /
// MakeItsLists / IsUnique / generate_mask (src/system.cpp:204-500,1718-1790),
// Solve_scf::CheckInput (src/solve_scf.cpp:114-281), Input parsing
// (src/input.cpp:3-78,286-429,706-740).
#pragma once
#include <string>
#include <vector>
#include <cstdint>

namespace nscf {

// Constant used by the reference for pi (src/namics.cpp:49).  It enters the
// curved-lattice weights and the Markov-2 bending weights, so the same value
// is used here.
constexpr double kPIE = 3.14159265;

enum Geometry : int { GEOM_PLANAR = 0, GEOM_CYLINDRICAL = 1, GEOM_SPHERICAL = 2 };
enum BoundaryKind : int { BC_MIRROR = 0, BC_PERIODIC = 1, BC_SURFACE = 2 };
enum LatticeType : int { LAT_SIMPLE_CUBIC = 0, LAT_HEXAGONAL = 1 };
// SEG_TAGGED: the segment sits on its mask only, with weight 1 there and no
// field (Segment::PrepareForCalculations src/segment.cpp:859-879).
enum SegFreedom : int { SEG_FREE = 0, SEG_FROZEN = 1, SEG_PINNED = 2, SEG_TAGGED = 3 };
// MOL_NEUTRALIZER: its bulk volume fraction follows from electroneutrality
// of the bulk at every evaluation (System::ComputePhis src/system.cpp:2487-2516).
// MOL_TAGGED: a molecule with a tagged segment; n is the number of tagged
// sites and its freedom key is ignored (Molecule::CheckInput
// src/molecule.cpp:455-459,578-580, System::CheckInput src/system.cpp:606-611).
enum MolFreedom : int { MOL_FREE = 0, MOL_RESTRICTED = 1, MOL_SOLVENT = 2, MOL_NEUTRALIZER = 3, MOL_TAGGED = 4 };
enum SolverMethod : int {
  SOLVER_DIIS = 0,
  SOLVER_PSEUDOHESSIAN = 1,
  SOLVER_HESSIAN = 2,
  SOLVER_LBFGS = 3,
  SOLVER_BRR = 4
};

struct LatticeSpec {
  int gradients = 1;
  int geometry = GEOM_PLANAR;
  int MX = 0, MY = 0, MZ = 0;   // interior layers (times fjc)
  int fjc = 1, FJC = 3;
  int lattice_type = LAT_SIMPLE_CUBIC;
  double lambda = 1.0 / 6.0;
  // bc[0..2] lower x,y,z; bc[3..5] upper x,y,z (reference BC[6] order)
  int bc[6] = {BC_MIRROR, BC_MIRROR, BC_MIRROR, BC_MIRROR, BC_MIRROR, BC_MIRROR};
  double offset_first_layer = 0.0;  // in grid units (times fjc)
  double bond_length = 0.0;         // `bondlength` (m), 0 when not given
  bool stencil_full = true;
  // derived
  int M = 0, JX = 0, JY = 0, JZ = 1;
  int BX1 = 0, BXM = 0, BY1 = 0, BYM = 0, BZ1 = 0, BZM = 0;
  double volume = 0.0;
  // Site-dependent weights (1 and 2 gradients).  L is the site volume
  // (1 inside a planar lattice, 0 on ghosts); lambda_1/lambda1 are the
  // weights toward the lower/upper x neighbour, lambda0 the self weight.
  std::vector<double> L, lambda_1, lambda0, lambda1;
  // Markov-2 reflection weights for curved 1-gradient lattices.
  std::vector<double> l1, l11, l_1, l_11;
  // Lattice refinement (fjc > 1, 1 or 2 gradients): ghost layers 0..fjc-1 and
  // MX+fjc..MX+2fjc-1 take their values from bx1f[k] / bxmf[k]; LAMBDA holds
  // FJC rows of M weights, row j (j < fjc) toward the site fjc-j below, row
  // FJC-1-j toward the site fjc-j above, row fjc the site itself.
  std::vector<int> bx1f, bxmf;
  std::vector<int> by1f, bymf;  // the same for y (2 and 3 gradients)
  std::vector<int> bz1f, bzmf;  // and for z (3 gradients)
  std::vector<double> LAMBDA;
};

// Largest refinement the device path handles (FJC_choices up to 9).
constexpr int kMaxFjcSpec = 4;

struct SegmentSpec {
  std::string name;
  int freedom = SEG_FREE;
  std::vector<int> mask;       // M entries, 1 on frozen / pinned sites
  std::vector<double> chi;     // n_seg entries, symmetric chi row
  int field_block = -1;        // which block of x holds this segment's field
  // Electrostatics (Segment::CheckInput src/segment.cpp:1056-1091): charge
  // per segment in units of e, relative dielectric constant, and a fixed
  // dimensionless surface potential e psi0 / kT on a frozen segment.
  double valence = 0.0, epsilon = 80.0;
  bool fixed_psi0 = false;
  double psi0 = 0.0;
};

struct BlockSpec {
  int seg = 0;
  int n = 0;
};

struct MoleculeSpec {
  std::string name;
  std::vector<BlockSpec> blocks;
  std::vector<int> molmon;     // distinct segments in order of appearance
  int chainlength = 0;
  int freedom = MOL_FREE;
  bool pinned = false;
  bool tagged = false;         // contains a tagged segment (one per molecule)
  bool ring = false;           // linear ring: first and last one-segment blocks merged
  bool has_n = false, has_theta = false;
  double phibulk = 0.0, n = 0.0, theta = 0.0;
  int markov = 1;
  double k_stiff = 0.0;
  bool save_memory = false;
  int size = 1;                // direction components per site (Markov 2: FJC in 1-D)
  std::vector<double> P;       // Markov-2 bending weights, FJC entries
  // index of the molmon entry for every block
  std::vector<int> block_molmon;
  // Branched architecture (square brackets in the composition,
  // Molecule::GenerateTree src/molecule.cpp:1222-1280): the generation of
  // every block, and per generation its first / last block and the contour
  // index of its first segment / one past its last (first_s, last_s).
  bool branched = false;
  std::vector<int> gnr, first_b, last_b, first_s, last_s;
  // Dendrimer (`@dend(A,(B)n,f; ...)`, Molecule::Decomposition
  // src/molecule.cpp:1437-1555): per generation a branch-point block and the
  // spacer blocks of each arm type, along one path from the core; block_deg
  // is each block's degeneracy (how many copies the molecule holds), n_arm
  // the number of arms of each arm type, and first_b / last_b / first_s /
  // last_s the spacer blocks and path indices of each arm type.  Linear and
  // branched molecules have block_deg 1.
  bool dendrimer = false;
  std::vector<int> block_deg, n_arm;
  // dendrimers: the arms of each generation (first_a / last_a index the
  // per-arm n_arm, first_b / last_b, first_s / last_s); a symmetric
  // dendrimer has one arm type per generation, an asymmetric one several
  std::vector<int> first_a, last_a;
  bool asymmetric = false;
  // Comb (`@comb((A)5; A,(B)3,(A)6,4; (A)7)`, Molecule::Decomposition
  // src/molecule.cpp:1556-1735): generation 0 the head of the backbone, 1
  // the tooth (degeneracy n_arm[0]), 2 .. n_arm[0]+1 the spacers, each after
  // a one-segment branch point, and n_arm[0]+2 the tail.
  bool comb = false;
  // propagated by a row schedule (build_branch_plan) instead of a linear chain
  bool scheduled() const { return branched || dendrimer || comb; }
};

// The propagator schedule of one branched molecule: a list of row
// operations over M-site rows.  Rows 0 .. chainlength-1 are the end-point
// distributions of the forward pass (one per contour index), rows
// chainlength and chainlength + 1 the two backward rows, the rest scratch
// rows of the junctions.  `seg` is a segment type; -1 in BR_PROP means no
// segment weight (the reference's UNITY).
enum BranchOpKind : int {
  BR_INIT = 0,      // R[a] = G1[seg]
  BR_PROP = 1,      // R[a] = W(R[b]) * G1[seg]
  BR_COPY = 2,      // R[a] = R[b]
  BR_TIMES = 3,     // R[a] = R[b] * R[c]
  BR_ADDPHI = 4,    // rho[c] += f * R[a] * R[b]   (c: molecule-segment index; f = 1 or a degeneracy)
  BR_ONE = 5,       // R[a] = 1
  BR_TIMES_G1 = 6,  // R[a] = R[b] * G1[seg]
  BR_DIV = 7,       // R[a] = R[b] != 0 ? R[a] / R[b] : 0
  BR_DIV_G1 = 8,    // R[a] = G1[seg] != 0 ? R[a] / G1[seg] : 0
};
struct BranchOp {
  int kind, a, b, c, seg;
  double f = 1.0;
};
struct BranchPlan {
  std::vector<BranchOp> forward, backward;
  int gn_row = 0;  // the row whose weighted sum is GN
  int rows = 0;    // rows the schedule uses
};
// Builds the schedule of molecule m (branched, Markov 1, no save_memory) by
// the reference's recursion (mol_branched::ForwardBra / BackwardBra
// src/mol_branched.cpp:166-300, Molecule::propagate_forward / _backward
// src/molecule.cpp:2300-2446, mol_branched::ComputePhi :303-420).  Returns
// false with a message when the composition leaves rows the reference would
// read before writing them.
bool build_branch_plan(const MoleculeSpec &m, BranchPlan &plan, std::string &err);

struct SolverSpec {
  int method = SOLVER_PSEUDOHESSIAN;
  double tolerance = 1e-7;
  double deltamax = 0.1;
  double deltamin = 0.0;       // smallest step bound of the LBFGS trust region
  int m = 10;                  // DIIS history, or the LBFGS correction pairs (default 6)
  int iterationlimit = 1000;
  bool max_g = false;          // stop on max|g| instead of ||g||_2
};

// One requested output quantity: `<type> : key : name : prop`, e.g.
// `kal : mol : pol : GN` (key sys / mol / mon / newton / lat).
struct OutputItem {
  std::string key, name, prop;
};

// Requests of one output type (Output::CheckInput / Load,
// src/output.cpp:32-178): present when an `output : <type> : ...` line exists.
struct OutputRequest {
  bool present = false;
  bool write = true;
  bool append = false;         // kal: rows of later starts are always appended
  bool append_given = false;
  bool write_bounds = false;   // pro: include the ghost layers
  std::string sep = "_";       // header_separator (`classic` gives ':')
  std::vector<OutputItem> items;
};

struct OutputSpec {
  OutputRequest kal, pro, vtk, vec, pos;
  bool use_output_folder = true;     // strip directories from the file stem
  std::string path = "./output/";    // `out_info : folder : path : ...`
};

// A `var : <kind>-<name> : scan : <property>` request (Variate::CheckInput,
// src/variate.cpp:64-345): the property is stepped from its value in the
// input to end_value, linearly by `step` or exponentially with `steps` per
// decade.
struct VarScan {
  bool present = false;
  std::string kind, name, prop;  // kind mol (theta, n, phibulk) or mon (chi_X)
  double start_value = 0, end_value = 0, step = 0;
  int steps = 1;
  bool exponential = false;
  int num_of_steps = 0;          // scan points after the first
};

// A `var : mol-<name> : search : theta | n | phibulk` request paired with a
// target, `var : sys-<name> : grand_potential | free_energy : value` or
// `var : mol-<name> : theta | n | phibulk | mu : value` (Variate::CheckInput,
// Molecule::PutVarInfo src/molecule.cpp:717-776, System::PutVarInfo
// src/system.cpp:1365-1400): the searched property is varied until the
// target's error (Molecule::GetError src/molecule.cpp:942-969,
// System::GetError src/system.cpp:1405-1427) is within super_tolerance.
struct VarSearch {
  bool present = false;
  std::string name, prop;           // searched molecule and property
  int mol = -1;
  double start_value = 0;
  std::string target_kind;          // "sys" or "mol"
  std::string target_prop;          // free_energy, grand_potential; theta, n, phibulk, mu
  int target_mol = -1;
  double target_value = 0;
  double super_tolerance = 0;       // newton super_tolerance (default 10 x tolerance)
  double super_deltamax = 0.5;      // newton super_deltamax
  int super_iterationlimit = 0;     // newton super_iterationlimit (default iterationlimit / 10)
};

struct SystemSpec {
  std::string sys_name = "NN";
  std::string newton_name = "NN";
  LatticeSpec lat;
  std::vector<SegmentSpec> seg;
  std::vector<MoleculeSpec> mol;
  SolverSpec solver;
  int solvent = -1;
  std::vector<int> itmon;      // segments owning a field block, in x order
  std::vector<int> ksam;       // M: 1 on accessible sites, 0 on frozen sites
  std::string initial_guess = "previous_result";
  std::string guess_inputfile;
  std::string final_guess = "next_problem";
  std::string guess_outputfile;
  int start = 1;               // the start block this system was built for
  OutputSpec out;
  VarScan scan;
  VarSearch search;
  std::string lattice_type_name, method_name = "pseudohessian", stop_criterion_name = "norm_of_g";
  std::vector<std::string> mol_composition, seg_freedom_name, mol_freedom_name;
  std::vector<double> seg_valence;
  // Electrostatics (System::CheckInput src/system.cpp:701-729): on when a
  // molecule carries a net charge; then x ends with one more block of M, the
  // dimensionless potential psi, and g with its Poisson residual.
  bool charged = false;
  int neutralizer = -1;        // the molecule with freedom neutralizer, or -1
  bool grad_epsilon = false;   // segments differ in epsilon
  bool fixed_psi0 = false;     // a frozen segment fixes the potential
  std::vector<double> mol_charge;  // Molecule::Charge: net valence per segment
  int nblocks_x() const { return (int)itmon.size() + (charged ? 1 : 0); }
  int nvar() const { return nblocks_x() * lat.M; }
  int max_chain() const {
    int n = 0;
    for (auto &m : mol) n = m.chainlength > n ? m.chainlength : n;
    return n;
  }
};

// Parse the reference's `.in` text and build the system for the given start
// block (1-based).  On failure returns false and fills err.
bool parse_input(const std::string &text, int start, SystemSpec &out, std::string &err);

// Number of `start` blocks in the text.
int count_starts(const std::string &text);

// The text with `line` inserted just before the start-th `start` line, so it
// overrides the settings of that block only from there on.
std::string insert_before_start(const std::string &text, int start, const std::string &line);

// Value of the scanned property at scan point k (Molecule::UpdateVarInfo,
// Segment::UpdateVarInfo, src/molecule.cpp:884-934, src/segment.cpp:1440-1480).
double scan_value(const VarScan &v, int k);

// Read a reference guess file (Lattice::ReadGuess format, src/lattice.cpp:926-968)
// and map its field blocks onto x by segment name (Solve_scf::Guess,
// src/solve_scf.cpp:489-524).  Returns false if the file does not fit.
bool read_guess_text(const std::string &text, const SystemSpec &s, std::vector<double> &x,
                     std::string &err);

// JSON description of the parsed system (names, sizes, field blocks), used by
// host tools and tests to lay out fields and results.
std::string describe_json(const SystemSpec &s);

// Lattice helpers shared by host code.
inline int site(const LatticeSpec &l, int x, int y = 0, int z = 0) {
  return x * l.JX + y * l.JY + z * l.JZ;
}

}  // namespace nscf
