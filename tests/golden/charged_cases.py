"""Charged systems (electrostatics on 1-gradient lattices) for reference
fixtures and parity tests.

Each case is `.in` text in the reference's input format; `eval` lists
(seed, amplitude) pairs of uniform random fields (the potential block
included) at which the residual is evaluated, `converged` the DIIS runs.
The cases cover: a charged polymer adsorbing on a wall with a neutralizer
counterion and free salt; a charged brush (pinned first segment) in salt
with segments of different dielectric constant; cylindrical and spherical
lattices with an offset; stiff (Markov 2) charged chains between two
walls, against one wall with a mirrored far boundary, and on a refined
hexagonal lattice; a refined planar hexagonal lattice (fjc 2); frozen plates that fix the potential
(`e.psi0/kT`, planar and cylindrical, the latter with segments of
different dielectric constant); 2-gradient cylindrical and planar lattices
and a 3-gradient box with frozen particles, without and with a fixed
potential on the particle.
"""

WALL = """
lat : flat : gradients : 1
lat : flat : n_layers : {n}
lat : flat : lattice_type : simple_cubic
lat : flat : lowerbound : surface
lat : flat : bondlength : 3e-10
mon : S : freedom : frozen
mon : S : frozen_range : lowerbound
"""

SALT = """
mon : Na : freedom : free
mon : Cl : freedom : free
mon : Na : valence : 1
mon : Cl : valence : -1
mon : W : freedom : free
mol : water : composition : (W)1
mol : water : freedom : solvent
"""

CHARGED = {
    # a weakly charged polymer adsorbing on an uncharged wall; Na
    # neutralizes the polymer, Cl is free salt
    "el1d_poly": dict(
        text=WALL.format(n=40) + SALT + """
mon : A : freedom : free
mon : A : valence : -0.2
mon : S : chi_A : -1.2
mon : A : chi_W : 0.5
mol : pol : composition : (A)50
mol : pol : freedom : restricted
mol : pol : theta : 3.0
mol : Na : composition : (Na)1
mol : Na : freedom : neutralizer
mol : Cl : composition : (Cl)1
mol : Cl : freedom : free
mol : Cl : phibulk : 0.01
""",
        eval=[(101, 0.3)],
        converged=[dict(method="DIIS", tol=1e-8)],
    ),
    # a charged brush: pinned first segment, no neutralizer (the free salt's
    # bulk is neutral), a low-epsilon polymer (grad_epsilon on)
    "el1d_brush": dict(
        text=WALL.format(n=30) + SALT + """
mon : X : freedom : pinned
mon : X : pinned_range : firstlayer
mon : A : freedom : free
mon : A : valence : -0.5
mon : X : valence : -0.5
mon : A : epsilon : 5
mon : X : epsilon : 5
mon : A : chi_W : 0.3
mol : brush : composition : (X)1(A)24
mol : brush : freedom : restricted
mol : brush : n : 0.05
mol : Na : composition : (Na)1
mol : Na : freedom : free
mol : Na : phibulk : 0.005
mol : Cl : composition : (Cl)1
mol : Cl : freedom : free
mol : Cl : phibulk : 0.005
""",
        eval=[(102, 0.2)],
        converged=[dict(method="DIIS", tol=1e-8)],
    ),
    # cylindrical lattice with an offset: a charged polymer around a rod
    "el1d_cyl": dict(
        text="""
lat : flat : gradients : 1
lat : flat : geometry : cylindrical
lat : flat : n_layers : 30
lat : flat : offset_first_layer : 4
lat : flat : lattice_type : simple_cubic
lat : flat : lowerbound : surface
lat : flat : bondlength : 5e-10
mon : S : freedom : frozen
mon : S : frozen_range : lowerbound
""" + SALT + """
mon : A : freedom : free
mon : A : valence : -0.3
mon : S : chi_A : -1.0
mol : pol : composition : (A)40
mol : pol : freedom : restricted
mol : pol : theta : 60
mol : Na : composition : (Na)1
mol : Na : freedom : neutralizer
mol : Cl : composition : (Cl)1
mol : Cl : freedom : free
mol : Cl : phibulk : 0.02
""",
        eval=[(103, 0.2)],
        converged=[dict(method="DIIS", tol=1e-8)],
    ),
    # spherical lattice (the Jacobi form of the potential sweep)
    "el1d_sph": dict(
        text="""
lat : flat : gradients : 1
lat : flat : geometry : spherical
lat : flat : n_layers : 25
lat : flat : offset_first_layer : 3
lat : flat : lattice_type : simple_cubic
lat : flat : lowerbound : surface
lat : flat : bondlength : 5e-10
mon : S : freedom : frozen
mon : S : frozen_range : lowerbound
""" + SALT + """
mon : A : freedom : free
mon : A : valence : -0.2
mon : S : chi_A : -1.0
mol : pol : composition : (A)20
mol : pol : freedom : restricted
mol : pol : theta : 200
mol : Na : composition : (Na)1
mol : Na : freedom : neutralizer
mol : Cl : composition : (Cl)1
mol : Cl : freedom : free
mol : Cl : phibulk : 0.02
""",
        eval=[(104, 0.2)],
        converged=[dict(method="DIIS", tol=1e-8)],
    ),
    # a stiff charged chain (Markov 2) in a planar gap between two walls
    "el1d_stiff": dict(
        text="""
lat : flat : gradients : 1
lat : flat : n_layers : 30
lat : flat : lattice_type : simple_cubic
lat : flat : lowerbound : surface
lat : flat : upperbound : surface
lat : flat : bondlength : 3e-10
mon : S : freedom : frozen
mon : S : frozen_range : lowerbound
mon : T : freedom : frozen
mon : T : frozen_range : upperbound
""" + SALT + """
mon : A : freedom : free
mon : A : valence : -0.4
mon : S : chi_A : -0.8
mon : T : chi_A : -0.8
mol : pol : composition : (A)30
mol : pol : freedom : restricted
mol : pol : theta : 2.0
mol : pol : Markov : 2
mol : pol : k_stiff : 1
mol : Na : composition : (Na)1
mol : Na : freedom : neutralizer
mol : Cl : composition : (Cl)1
mol : Cl : freedom : free
mol : Cl : phibulk : 0.01
""",
        eval=[(105, 0.2)],
        converged=[dict(method="DIIS", tol=1e-8)],
    ),
    # refined planar hexagonal lattice (two grid layers per bond)
    "el1d_fjc2": dict(
        text="""
lat : flat : gradients : 1
lat : flat : n_layers : 20
lat : flat : FJC_choices : 5
lat : flat : lattice_type : hexagonal
lat : flat : lowerbound : surface
lat : flat : bondlength : 3e-10
mon : S : freedom : frozen
mon : S : frozen_range : lowerbound
""" + SALT + """
mon : A : freedom : free
mon : A : valence : -0.2
mon : S : chi_A : -1.0
mol : pol : composition : (A)20
mol : pol : freedom : restricted
mol : pol : theta : 1.5
mol : Na : composition : (Na)1
mol : Na : freedom : neutralizer
mol : Cl : composition : (Cl)1
mol : Cl : freedom : free
mol : Cl : phibulk : 0.01
""",
        eval=[(106, 0.2)],
        converged=[dict(method="DIIS", tol=1e-8)],
    ),
    # stiff charged chains against a wall with the far boundary mirrored:
    # the ghosts there carry the reference's paired-component end points
    "el1d_stiff_mirror": dict(
        text=WALL.format(n=30) + SALT + """
mon : A : freedom : free
mon : A : valence : -0.3
mon : S : chi_A : -0.6
mol : pol : composition : (A)25
mol : pol : freedom : restricted
mol : pol : theta : 3.0
mol : pol : Markov : 2
mol : pol : k_stiff : 2
mol : Na : composition : (Na)1
mol : Na : freedom : neutralizer
mol : Cl : composition : (Cl)1
mol : Cl : freedom : free
mol : Cl : phibulk : 0.01
""",
        eval=[(109, 0.2)],
        converged=[dict(method="DIIS", tol=1e-8)],
    ),
    # the same on a refined hexagonal lattice (FJC direction components)
    "el1d_hex_m2": dict(
        text="""
lat : flat : gradients : 1
lat : flat : n_layers : 20
lat : flat : FJC_choices : 5
lat : flat : lattice_type : hexagonal
lat : flat : lowerbound : surface
lat : flat : bondlength : 3e-10
mon : S : freedom : frozen
mon : S : frozen_range : lowerbound
""" + SALT + """
mon : A : freedom : free
mon : A : valence : -0.2
mon : S : chi_A : -0.8
mol : pol : composition : (A)20
mol : pol : freedom : restricted
mol : pol : theta : 1.5
mol : pol : Markov : 2
mol : pol : k_stiff : 1
mol : Na : composition : (Na)1
mol : Na : freedom : neutralizer
mol : Cl : composition : (Cl)1
mol : Cl : freedom : free
mol : Cl : phibulk : 0.01
""",
        eval=[(110, 0.2)],
        converged=[dict(method="DIIS", tol=1e-8)],
    ),
    # 2-gradient cylindrical lattice: a charged polymer around a frozen rod
    # (the Jacobi potential sweep of LGrad2)
    "el2d_cyl": dict(
        text="""
lat : cyl : gradients : 2
lat : cyl : geometry : cylindrical
lat : cyl : n_layers_x : 12
lat : cyl : n_layers_y : 10
lat : cyl : lattice_type : simple_cubic
lat : cyl : bondlength : 5e-10
mon : S : freedom : frozen
mon : S : frozen_range : 1,1;2,10
""" + SALT + """
mon : A : freedom : free
mon : A : valence : -0.3
mon : S : chi_A : -1.0
mol : pol : composition : (A)15
mol : pol : freedom : restricted
mol : pol : theta : 40
mol : Na : composition : (Na)1
mol : Na : freedom : neutralizer
mol : Cl : composition : (Cl)1
mol : Cl : freedom : free
mol : Cl : phibulk : 0.02
""",
        eval=[(111, 0.2)],
        converged=[dict(method="DIIS", tol=1e-8)],
    ),
    # 2-gradient planar lattice with a frozen particle, periodic in y, and a
    # low-epsilon polymer (the upwind sweep of LG2Planar in x and y)
    "el2d_planar": dict(
        text="""
lat : L : gradients : 2
lat : L : n_layers_x : 10
lat : L : n_layers_y : 12
lat : L : lattice_type : simple_cubic
lat : L : lowerbound_y : periodic
lat : L : upperbound_y : periodic
lat : L : bondlength : 3e-10
mon : S : freedom : frozen
mon : S : frozen_range : 4,5;6,7
""" + SALT + """
mon : A : freedom : free
mon : A : valence : -0.4
mon : A : epsilon : 20
mon : S : chi_A : -1.2
mol : pol : composition : (A)12
mol : pol : freedom : restricted
mol : pol : theta : 6
mol : Na : composition : (Na)1
mol : Na : freedom : neutralizer
mol : Cl : composition : (Cl)1
mol : Cl : freedom : free
mol : Cl : phibulk : 0.01
""",
        eval=[(112, 0.2)],
        converged=[dict(method="DIIS", tol=1e-8)],
    ),
    # 3-gradient box with a frozen particle (the symmetric Gauss-Seidel
    # sweep of LGrad3)
    "el3d_box": dict(
        text="""
lat : box : gradients : 3
lat : box : n_layers_x : 8
lat : box : n_layers_y : 7
lat : box : n_layers_z : 9
lat : box : lattice_type : simple_cubic
lat : box : bondlength : 3e-10
mon : S : freedom : frozen
mon : S : frozen_range : 3,3,4;4,5,5
""" + SALT + """
mon : A : freedom : free
mon : A : valence : -0.3
mon : S : chi_A : -1.0
mol : pol : composition : (A)10
mol : pol : freedom : restricted
mol : pol : theta : 12
mol : Na : composition : (Na)1
mol : Na : freedom : neutralizer
mol : Cl : composition : (Cl)1
mol : Cl : freedom : free
mol : Cl : phibulk : 0.01
""",
        eval=[(113, 0.2)],
        converged=[dict(method="DIIS", tol=1e-8)],
    ),
    # fixed potentials on 2 and 3 gradients (evaluations: the reference's
    # fixed-potential sweeps are the same overwrite as on 1 gradient)
    "el2d_fixed_planar": dict(
        text="""
lat : L : gradients : 2
lat : L : n_layers_x : 10
lat : L : n_layers_y : 12
lat : L : lattice_type : simple_cubic
lat : L : lowerbound_y : periodic
lat : L : upperbound_y : periodic
lat : L : bondlength : 3e-10
mon : P : freedom : frozen
mon : P : frozen_range : 4,5;6,7
mon : P : e.psi0/kT : 1.0
""" + SALT + """
mon : A : freedom : free
mon : A : valence : -0.3
mon : A : epsilon : 20
mon : P : chi_A : -0.5
mol : pol : composition : (A)12
mol : pol : freedom : restricted
mol : pol : theta : 6
mol : Na : composition : (Na)1
mol : Na : freedom : neutralizer
mol : Cl : composition : (Cl)1
mol : Cl : freedom : free
mol : Cl : phibulk : 0.01
""",
        eval=[(114, 0.2)],
        converged=[],
    ),
    "el2d_fixed_cyl": dict(
        text="""
lat : cyl : gradients : 2
lat : cyl : geometry : cylindrical
lat : cyl : n_layers_x : 12
lat : cyl : n_layers_y : 10
lat : cyl : lattice_type : simple_cubic
lat : cyl : bondlength : 5e-10
mon : P : freedom : frozen
mon : P : frozen_range : 1,1;2,10
mon : P : e.psi0/kT : -1.5
""" + SALT + """
mon : A : freedom : free
mon : A : valence : 0.3
mon : A : epsilon : 10
mon : P : chi_A : -0.5
mol : pol : composition : (A)15
mol : pol : freedom : restricted
mol : pol : theta : 40
mol : Na : composition : (Na)1
mol : Na : freedom : free
mol : Na : phibulk : 0.02
mol : Cl : composition : (Cl)1
mol : Cl : freedom : neutralizer
""",
        eval=[(115, 0.2)],
        converged=[],
    ),
    "el3d_fixed": dict(
        text="""
lat : box : gradients : 3
lat : box : n_layers_x : 8
lat : box : n_layers_y : 7
lat : box : n_layers_z : 9
lat : box : lattice_type : simple_cubic
lat : box : bondlength : 3e-10
mon : P : freedom : frozen
mon : P : frozen_range : 3,3,4;4,5,5
mon : P : e.psi0/kT : 1.0
""" + SALT + """
mon : A : freedom : free
mon : A : valence : -0.3
mon : A : epsilon : 15
mon : P : chi_A : -1.0
mol : pol : composition : (A)10
mol : pol : freedom : restricted
mol : pol : theta : 12
mol : Na : composition : (Na)1
mol : Na : freedom : neutralizer
mol : Cl : composition : (Cl)1
mol : Cl : freedom : free
mol : Cl : phibulk : 0.01
""",
        eval=[(116, 0.2)],
        converged=[],
    ),
    # a frozen plate at the first layer fixing the potential, a charged
    # polymer and salt in front of it
    "el1d_fixed": dict(
        text=WALL.format(n=30) + SALT + """
mon : P : freedom : frozen
mon : P : frozen_range : 1;1
mon : P : e.psi0/kT : 1.5
mon : A : freedom : free
mon : A : valence : -0.3
mon : P : chi_A : -0.5
mol : pol : composition : (A)20
mol : pol : freedom : restricted
mol : pol : theta : 1.5
mol : Na : composition : (Na)1
mol : Na : freedom : neutralizer
mol : Cl : composition : (Cl)1
mol : Cl : freedom : free
mol : Cl : phibulk : 0.01
""",
        eval=[(107, 0.2)],
        converged=[],
    ),
    # the same on a cylindrical lattice, with a low-epsilon polymer
    "el1d_fixed_cyl": dict(
        text="""
lat : flat : gradients : 1
lat : flat : geometry : cylindrical
lat : flat : n_layers : 30
lat : flat : offset_first_layer : 2
lat : flat : lattice_type : simple_cubic
lat : flat : lowerbound : surface
lat : flat : bondlength : 5e-10
mon : S : freedom : frozen
mon : S : frozen_range : lowerbound
mon : P : freedom : frozen
mon : P : frozen_range : 1;2
mon : P : e.psi0/kT : -1
""" + SALT + """
mon : A : freedom : free
mon : A : valence : 0.2
mon : A : epsilon : 10
mon : P : chi_A : -0.3
mol : pol : composition : (A)25
mol : pol : freedom : restricted
mol : pol : theta : 30
mol : Na : composition : (Na)1
mol : Na : freedom : free
mol : Na : phibulk : 0.01
mol : Cl : composition : (Cl)1
mol : Cl : freedom : neutralizer
""",
        eval=[(108, 0.2)],
        converged=[],
    ),
}
