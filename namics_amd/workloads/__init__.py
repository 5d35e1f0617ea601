"""Synthetic SCF workloads (the benchmark configurations), as `.in` text.

Each generator returns input text in the reference's format plus seeded
initial fields, so the same problem can be handed to the GPU library and to
the reference program.
"""
import numpy as np

PIE = 3.14159265


def brush_cylinder(chi, n_r=64, n_z=128, N=200, sigma=0.05, lattice="simple_cubic", tol=1e-7):
    """2-gradient cylindrical polymer brush on a rod: a frozen rod S fills the
    first radial layer (radius 1), chains (X)1(A)N-1 are pinned with their X
    end on the rod surface (second radial layer, every axial position),
    monomeric solvent W, chi(A,W) = chi.  Grafting density sigma is chains
    per unit rod area: n = sigma * 2 PIE R n_z with R = 1."""
    n = sigma * 2 * PIE * 1.0 * n_z
    text = """
lat : cyl : gradients : 2
lat : cyl : geometry : cylindrical
lat : cyl : n_layers_x : {nr}
lat : cyl : n_layers_y : {nz}
lat : cyl : lattice_type : {lat}
mon : S : freedom : frozen
mon : S : frozen_range : 1,1;1,{nz}
mon : X : freedom : pinned
mon : X : pinned_range : 2,1;2,{nz}
mon : A : freedom : free
mon : W : freedom : free
mon : A : chi_W : {chi}
mol : brush : composition : (X)1(A){na}
mol : brush : freedom : restricted
mol : brush : n : {n!r}
mol : water : composition : (W)1
mol : water : freedom : solvent
newton : isaac : method : DIIS
newton : isaac : tolerance : {tol!r}
newton : isaac : iterationlimit : 5000
""".format(nr=n_r, nz=n_z, lat=lattice, chi=chi, na=N - 1, n=n, tol=tol)
    return text


def diblock_box(n=96, frac=0.05, chi_bw=1.6, chi_ab=1.6, tol=1e-7, save_memory=True, itlim=5000):
    """3-gradient periodic simple-cubic box of n^3 sites: diblock (A)10(B)10,
    restricted to theta = frac * n^3 monomers, in monomeric solvent W, with
    chi(B,W) = chi(A,B) = chi_bw / chi_ab and chi(A,W) = 0."""
    theta = frac * n ** 3
    return """
lat : box : gradients : 3
lat : box : n_layers_x : {n}
lat : box : n_layers_y : {n}
lat : box : n_layers_z : {n}
lat : box : lattice_type : simple_cubic
lat : box : lowerbound_x : periodic
lat : box : upperbound_x : periodic
lat : box : lowerbound_y : periodic
lat : box : upperbound_y : periodic
lat : box : lowerbound_z : periodic
lat : box : upperbound_z : periodic
mon : A : freedom : free
mon : B : freedom : free
mon : W : freedom : free
mon : B : chi_W : {cbw!r}
mon : A : chi_B : {cab!r}
mol : surf : composition : (A)10(B)10
mol : surf : freedom : restricted
mol : surf : theta : {theta!r}
mol : surf : save_memory : {sm}
mol : water : composition : (W)1
mol : water : freedom : solvent
newton : isaac : method : DIIS
newton : isaac : tolerance : {tol!r}
newton : isaac : iterationlimit : {itlim}
""".format(n=n, cbw=chi_bw, cab=chi_ab, theta=theta, sm="true" if save_memory else "false", tol=tol,
           itlim=itlim)


def seeded_box_fields(names, n, seed, amp=0.1, well=1.0, width=4.0, well_seg="B"):
    """Initial fields for `diblock_box`: uniform noise in [-amp, amp] on every
    field block plus, on the block of `well_seg`, a Gaussian well of depth
    `well` and width `width` (sites) at the box centre, which nucleates one
    aggregate.  Layout: blocks in `names` order, each (n+2)^3 x-major with
    ghost layers (ghost values are never read)."""
    m = n + 2
    x = uniform_fields(len(names) * m ** 3, seed, amp).reshape(len(names), m, m, m)
    if well_seg in names:
        c = (n + 1) / 2.0
        r = np.arange(m) - c
        r2 = r[:, None, None] ** 2 + r[None, :, None] ** 2 + r[None, None, :] ** 2
        x[names.index(well_seg)] -= well * np.exp(-r2 / (2.0 * width ** 2))
    return x.ravel()


def uniform_fields(nvar, seed, amp):
    return np.random.default_rng(seed).uniform(-amp, amp, nvar)


def homopolymer_adsorption(layers, N, chi, chi_s, phib, markov=1, k_stiff=1.0, tol=1e-7):
    """1-gradient planar adsorption: frozen surface S on the lower ghost layer,
    homopolymer (A)N free at phib, monomeric solvent W."""
    extra = ""
    if markov == 2:
        extra = "mol : pol : Markov : 2\nmol : pol : k_stiff : %r\n" % k_stiff
    return """
lat : flat : gradients : 1
lat : flat : n_layers : {L}
lat : flat : lattice_type : simple_cubic
lat : flat : lowerbound : surface
mon : S : freedom : frozen
mon : S : frozen_range : lowerbound
mon : A : freedom : free
mon : W : freedom : free
mon : S : chi_A : {cs!r}
mon : A : chi_W : {c!r}
mol : pol : composition : (A){N}
mol : pol : freedom : free
mol : pol : phibulk : {pb!r}
{extra}mol : water : composition : (W)1
mol : water : freedom : solvent
newton : isaac : method : DIIS
newton : isaac : tolerance : {tol!r}
newton : isaac : iterationlimit : 5000
""".format(L=layers, cs=chi_s, c=chi, N=N, pb=phib, extra=extra, tol=tol)


def polyelectrolyte_adsorption(layers, N, valence, chi_s, salt, theta, markov=1, k_stiff=1.0, tol=1e-7):
    """1-gradient planar adsorption of a charged homopolymer: frozen
    uncharged surface S on the lower ghost layer, (A)N restricted at
    `theta` with `valence` per segment, counterions Na (+1) as the
    neutralizer, salt Cl (-1) free at volume fraction `salt`, monomeric
    solvent W, bond length 3e-10 m."""
    extra = ""
    if markov == 2:
        extra = "mol : pol : Markov : 2\nmol : pol : k_stiff : %r\n" % k_stiff
    return """
lat : flat : gradients : 1
lat : flat : n_layers : {L}
lat : flat : lattice_type : simple_cubic
lat : flat : lowerbound : surface
lat : flat : bondlength : 3e-10
mon : S : freedom : frozen
mon : S : frozen_range : lowerbound
mon : A : freedom : free
mon : W : freedom : free
mon : Na : freedom : free
mon : Cl : freedom : free
mon : A : valence : {v!r}
mon : Na : valence : 1
mon : Cl : valence : -1
mon : S : chi_A : {cs!r}
mol : pol : composition : (A){N}
mol : pol : freedom : restricted
mol : pol : theta : {th!r}
{extra}mol : Na : composition : (Na)1
mol : Na : freedom : neutralizer
mol : Cl : composition : (Cl)1
mol : Cl : freedom : free
mol : Cl : phibulk : {salt!r}
mol : water : composition : (W)1
mol : water : freedom : solvent
newton : isaac : method : DIIS
newton : isaac : tolerance : {tol!r}
newton : isaac : iterationlimit : 5000
""".format(L=layers, v=valence, cs=chi_s, N=N, th=theta, salt=salt, extra=extra, tol=tol)


def charged_draws(n, seed=2024):
    """Charged 1-gradient draws: layers U{30..200}, N log-uniform in
    [20, 500], valence U[-1, -0.05], chi_S U[-1.5, 0], salt log-uniform in
    [1e-3, 1e-1], adsorbed amount theta U[0.5, 5]; odd draws second-order
    Markov with k_stiff = 1."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        out.append(dict(kind="charged", layers=int(rng.integers(30, 201)),
                        N=int(round(float(np.exp(rng.uniform(np.log(20), np.log(500)))))),
                        valence=float(rng.uniform(-1.0, -0.05)), chi_s=float(rng.uniform(-1.5, 0.0)),
                        salt=float(np.exp(rng.uniform(np.log(1e-3), np.log(1e-1)))),
                        theta=float(rng.uniform(0.5, 5.0)), markov=2 if i % 2 == 1 else 1))
    return out


def batch_draws(n, seed=12345):
    """Configuration-4 style draws: layers U{50..400}, N log-uniform in
    [20,2000], chi U[0.3,1.2], chi_S U[-1.5,0], phib log-uniform in
    [1e-4,1e-1]; odd draws are second-order Markov with k_stiff = 1."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        L = int(rng.integers(50, 401))
        N = int(round(float(np.exp(rng.uniform(np.log(20), np.log(2000))))))
        chi = float(rng.uniform(0.3, 1.2))
        chi_s = float(rng.uniform(-1.5, 0.0))
        phib = float(np.exp(rng.uniform(np.log(1e-4), np.log(1e-1))))
        markov = 2 if i % 2 == 1 else 1
        out.append(dict(layers=L, N=N, chi=chi, chi_s=chi_s, phib=phib, markov=markov))
    return out


def carry_fields(x_prev, names_prev, names_new, M):
    """Start a new solve from a previous result the way the reference chains
    `start` blocks (Solve_scf::Guess, src/solve_scf.cpp:489-524): field blocks
    are copied by segment name; blocks without a match start at zero."""
    x = np.zeros(len(names_new) * M)
    for j, n in enumerate(names_new):
        if n in names_prev:
            i = names_prev.index(n)
            x[j * M:(j + 1) * M] = x_prev[i * M:(i + 1) * M]
    return x


def brush_job_input(guess_file, chis=(0.0, 0.5, 0.8), **kw):
    """One benchmark brush job as a reference input: one `start` block per chi
    value, the first from the fields in `guess_file` (guess-file format), each
    later one from the previous result, with a kal row per solve (iterations,
    residual, free energy, grand potential) appended to output/<name>.kal."""
    text = brush_cylinder(chis[0], **kw)
    text += ("sys : NN : initial_guess : file\nsys : NN : guess_inputfile : %s\n"
             "newton : isaac : e_info : false\n"
             "output : kal : append : true\nkal : newton : isaac : iterations\n"
             "kal : newton : isaac : residual\nkal : sys : NN : free_energy\n"
             "kal : sys : NN : grand_potential\nstart\n") % guess_file
    for c in chis[1:]:
        text += "mon : A : chi_W : %r\nsys : NN : initial_guess : previous_result\nstart\n" % c
    return text


def read_job_kal(path):
    """Rows of a kal file written by `brush_job_input`, one dict per solve:
    iterations, residual, free_energy, grand_potential."""
    with open(path) as f:
        lines = [l.rstrip("\n").split("\t") for l in f if l.strip()]
    keys = {"newton_isaac_iterations": "iterations", "newton_isaac_residual": "residual",
            "sys_NN_free_energy": "free_energy", "sys_NN_grand_potential": "grand_potential"}
    out = []
    for row in lines[1:]:
        r = {}
        for k, v in zip(lines[0], row):
            if k in keys:
                r[keys[k]] = float(v)
        r["iterations"] = int(r["iterations"])
        out.append(r)
    return out


def brush_chain_input(chis=(0.0, 0.5, 0.8), **kw):
    """The brush solves as one reference input with one `start` block per
    chi value (each solve starts from the previous result)."""
    text = brush_cylinder(chis[0], **kw) + "start\n"
    for c in chis[1:]:
        text += "mon : A : chi_W : %r\nstart\n" % c
    return text


def write_guess(path, names, dims, x, method="DIIS", fjc=1, charged=False):
    """Write fields in the reference's guess-file format (Lattice::StoreGuess,
    src/lattice.cpp:971-1019; read back by Lattice::ReadGuess) at full
    precision: method line, sizes (grid layers) and refinement, charge flag,
    field-block names, values (with charged=True, x ends with the potential
    block)."""
    mx, my, mz = dims
    with open(path, "w") as f:
        f.write("%s\n" % method)
        f.write(" %d\t%d\t%d\t%d\n" % (mx, my, mz, fjc))
        f.write("true\n" if charged else "false\n")
        f.write("%d\n" % len(names))
        for n in names:
            f.write("%s\n" % n)
        f.write("0\n")
        for v in x:
            f.write("%.17e\n" % v)
