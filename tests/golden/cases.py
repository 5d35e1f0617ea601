"""Small SCF systems used for golden fixtures and parity tests.

Each case is `.in` text in the reference's input format (no output, newton
method or guess lines: the fixture generator and the tests add those).
`eval` lists (seed, amplitude) pairs of uniform random fields at which the
residual is evaluated; `converged` lists solver runs.
"""

CFG1 = """
lat : flat : gradients : 1
lat : flat : n_layers : 100
lat : flat : lattice_type : simple_cubic
lat : flat : lowerbound : surface
mon : S : freedom : frozen
mon : S : frozen_range : lowerbound
mon : A : freedom : free
mon : W : freedom : free
mon : S : chi_A : -1.0
mon : A : chi_W : 0.5
mol : pol : composition : (A)100
mol : pol : freedom : free
mol : pol : phibulk : 0.01
mol : water : composition : (W)1
mol : water : freedom : solvent
newton : isaac : deltamax : 0.1
"""

CASES = {
    # configuration 1: planar homopolymer adsorption
    "cfg1_planar_adsorption": dict(
        text=CFG1,
        eval=[(1, 0.3), (2, 1.0)],
        converged=[dict(method="pseudohessian", tol=1e-12),
                   dict(method="DIIS", tol=1e-7)],
    ),
    # second-order Markov (bending energy k_stiff = 1) on the same lattice
    "planar_markov2": dict(
        text="""
lat : flat : gradients : 1
lat : flat : n_layers : 80
lat : flat : lattice_type : simple_cubic
lat : flat : lowerbound : surface
mon : S : freedom : frozen
mon : S : frozen_range : lowerbound
mon : A : freedom : free
mon : W : freedom : free
mon : S : chi_A : -1.2
mon : A : chi_W : 0.6
mol : pol : composition : (A)60
mol : pol : freedom : free
mol : pol : phibulk : 0.005
mol : pol : Markov : 2
mol : pol : k_stiff : 1
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(3, 0.3)],
        converged=[dict(method="DIIS", tol=1e-7)],
    ),
    # curved 1-gradient: cylindrical micelle of a restricted diblock
    "cyl1d_micelle": dict(
        text="""
lat : L : gradients : 1
lat : L : n_layers : 40
lat : L : geometry : cylindrical
lat : L : lattice_type : simple_cubic
mon : A : freedom : free
mon : B : freedom : free
mon : W : freedom : free
mon : B : chi_W : 1.5
mon : A : chi_B : 1.0
mol : surf : composition : (A)6(B)6
mol : surf : freedom : restricted
mol : surf : theta : 300
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(4, 0.5)],
    ),
    # curved 1-gradient: spherical particle with an adsorbing surface
    "sph1d_particle": dict(
        text="""
lat : L : gradients : 1
lat : L : n_layers : 50
lat : L : geometry : spherical
lat : L : offset_first_layer : 5
lat : L : lattice_type : simple_cubic
lat : L : lowerbound : surface
mon : S : freedom : frozen
mon : S : frozen_range : lowerbound
mon : A : freedom : free
mon : W : freedom : free
mon : S : chi_A : -1.5
mon : A : chi_W : 0.4
mol : pol : composition : (A)30
mol : pol : freedom : free
mol : pol : phibulk : 0.02
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(5, 0.3)],
    ),
    # hexagonal lattice, periodic boundaries, restricted homopolymer
    "hex_periodic_restricted": dict(
        text="""
lat : L : gradients : 1
lat : L : n_layers : 60
lat : L : lattice_type : hexagonal
lat : L : lowerbound : periodic
lat : L : upperbound : periodic
mon : A : freedom : free
mon : W : freedom : free
mon : A : chi_W : 0.7
mol : pol : composition : (A)20
mol : pol : freedom : restricted
mol : pol : theta : 8
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(6, 0.4)],
    ),
    # lattice refinement (FJC_choices 5: two grid layers per bond length),
    # planar adsorption next to a surface
    "fjc5_planar_adsorption": dict(
        text="""
lat : flat : gradients : 1
lat : flat : n_layers : 40
lat : flat : lattice_type : hexagonal
lat : flat : FJC_choices : 5
lat : flat : lowerbound : surface
mon : S : freedom : frozen
mon : S : frozen_range : lowerbound
mon : A : freedom : free
mon : W : freedom : free
mon : S : chi_A : -1.0
mon : A : chi_W : 0.5
mol : pol : composition : (A)50
mol : pol : freedom : free
mol : pol : phibulk : 0.01
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(31, 0.3), (32, 0.8)],
        converged=[dict(method="DIIS", tol=1e-10, seed=33, amp=0.05)],
    ),
    # refinement through b/l (three grid layers per bond), grafted brush on a
    # mirror, hexagonal lattice
    "fjc7_planar_brush": dict(
        text="""
lat : flat : gradients : 1
lat : flat : n_layers : 30
lat : flat : lattice_type : hexagonal
lat : flat : b/l : 3
mon : X : freedom : pinned
mon : X : pinned_range : firstlayer
mon : A : freedom : free
mon : W : freedom : free
mon : A : chi_W : 0.4
mol : brush : composition : (X)1(A)40
mol : brush : freedom : restricted
mol : brush : theta : 4.1
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(34, 0.3)],
        converged=[dict(method="DIIS", tol=1e-10, seed=35, amp=0.05)],
    ),
    # refined cylindrical micelle of a restricted diblock (refinement needs
    # the hexagonal lattice)
    "fjc5_cyl_micelle": dict(
        text="""
lat : L : gradients : 1
lat : L : n_layers : 25
lat : L : geometry : cylindrical
lat : L : lattice_type : hexagonal
lat : L : FJC_choices : 5
mon : A : freedom : free
mon : B : freedom : free
mon : W : freedom : free
mon : B : chi_W : 1.5
mon : A : chi_B : 1.0
mol : surf : composition : (A)6(B)6
mol : surf : freedom : restricted
mol : surf : theta : 150
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(36, 0.5)],
    ),
    # refined spherical particle with an offset and an adsorbing surface
    "fjc5_sph_particle": dict(
        text="""
lat : L : gradients : 1
lat : L : n_layers : 30
lat : L : geometry : spherical
lat : L : offset_first_layer : 4
lat : L : lattice_type : hexagonal
lat : L : FJC_choices : 5
lat : L : lowerbound : surface
mon : S : freedom : frozen
mon : S : frozen_range : lowerbound
mon : A : freedom : free
mon : W : freedom : free
mon : S : chi_A : -1.5
mon : A : chi_W : 0.4
mol : pol : composition : (A)30
mol : pol : freedom : free
mol : pol : phibulk : 0.02
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(37, 0.3)],
    ),
    # refined periodic lattice (FJC_choices 9), restricted homopolymer
    "fjc9_periodic_restricted": dict(
        text="""
lat : L : gradients : 1
lat : L : n_layers : 24
lat : L : lattice_type : hexagonal
lat : L : FJC_choices : 9
lat : L : lowerbound : periodic
lat : L : upperbound : periodic
mon : A : freedom : free
mon : W : freedom : free
mon : A : chi_W : 0.7
mol : pol : composition : (A)20
mol : pol : freedom : restricted
mol : pol : theta : 6
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(38, 0.4)],
    ),
    # second-order Markov on the hexagonal lattice (the general direction
    # loop with FJC = 3 components)
    "hex_markov2_planar": dict(
        text="""
lat : flat : gradients : 1
lat : flat : n_layers : 50
lat : flat : lattice_type : hexagonal
lat : flat : lowerbound : surface
mon : S : freedom : frozen
mon : S : frozen_range : lowerbound
mon : A : freedom : free
mon : W : freedom : free
mon : S : chi_A : -1.2
mon : A : chi_W : 0.5
mol : pol : composition : (A)40
mol : pol : freedom : free
mol : pol : phibulk : 0.005
mol : pol : Markov : 2
mol : pol : k_stiff : 1
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(41, 0.3)],
        converged=[dict(method="DIIS", tol=1e-10, seed=42, amp=0.05)],
    ),
    # second-order Markov on a refined hexagonal lattice (FJC = 5
    # components), adsorption next to a surface
    "fjc5_markov2_adsorption": dict(
        text="""
lat : flat : gradients : 1
lat : flat : n_layers : 30
lat : flat : lattice_type : hexagonal
lat : flat : FJC_choices : 5
lat : flat : lowerbound : surface
mon : S : freedom : frozen
mon : S : frozen_range : lowerbound
mon : A : freedom : free
mon : W : freedom : free
mon : S : chi_A : -1.0
mon : A : chi_W : 0.6
mol : pol : composition : (A)30
mol : pol : freedom : free
mol : pol : phibulk : 0.01
mol : pol : Markov : 2
mol : pol : k_stiff : 2
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(43, 0.3), (44, 0.6)],
        converged=[dict(method="DIIS", tol=1e-10, seed=45, amp=0.05)],
    ),
    # stiff grafted brush on a refined lattice (b/l 3, FJC = 7 components),
    # mirror boundaries
    "fjc7_markov2_brush": dict(
        text="""
lat : flat : gradients : 1
lat : flat : n_layers : 25
lat : flat : lattice_type : hexagonal
lat : flat : b/l : 3
mon : X : freedom : pinned
mon : X : pinned_range : firstlayer
mon : A : freedom : free
mon : W : freedom : free
mon : A : chi_W : 0.3
mol : brush : composition : (X)1(A)30
mol : brush : freedom : restricted
mol : brush : theta : 3.1
mol : brush : Markov : 2
mol : brush : k_stiff : 0.5
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(46, 0.3)],
        converged=[dict(method="DIIS", tol=1e-9, seed=47, amp=0.05)],
    ),
    # second-order Markov on a curved 1-gradient lattice
    "cyl1d_markov2": dict(
        text="""
lat : L : gradients : 1
lat : L : n_layers : 40
lat : L : geometry : cylindrical
lat : L : offset_first_layer : 3
lat : L : lattice_type : simple_cubic
mon : A : freedom : free
mon : W : freedom : free
mon : A : chi_W : 0.8
mol : pol : composition : (A)25
mol : pol : freedom : restricted
mol : pol : theta : 40
mol : pol : Markov : 2
mol : pol : k_stiff : 2
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(7, 0.3)],
    ),
    # second-order Markov on curved hexagonal lattices (reflection-weighted
    # steps off the axis), cylindrical and spherical
    "hex_markov2_cyl": dict(
        text="""
lat : L : gradients : 1
lat : L : n_layers : 30
lat : L : geometry : cylindrical
lat : L : offset_first_layer : 2
lat : L : lattice_type : hexagonal
lat : L : lowerbound : surface
mon : S : freedom : frozen
mon : S : frozen_range : lowerbound
mon : A : freedom : free
mon : W : freedom : free
mon : S : chi_A : -1.0
mon : A : chi_W : 0.6
mol : pol : composition : (A)20
mol : pol : freedom : restricted
mol : pol : theta : 30
mol : pol : Markov : 2
mol : pol : k_stiff : 1.5
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(51, 0.3)],
        converged=[dict(method="DIIS", tol=1e-9)],
    ),
    "hex_markov2_sph": dict(
        text="""
lat : L : gradients : 1
lat : L : n_layers : 25
lat : L : geometry : spherical
lat : L : offset_first_layer : 3
lat : L : lattice_type : hexagonal
mon : A : freedom : free
mon : W : freedom : free
mon : A : chi_W : 0.9
mol : pol : composition : (A)15
mol : pol : freedom : restricted
mol : pol : theta : 300
mol : pol : Markov : 2
mol : pol : k_stiff : 1
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(52, 0.3)],
        converged=[dict(method="DIIS", tol=1e-9, seed=53, amp=0.1)],
    ),
    # 2-gradient cylindrical polymer brush (small version of configuration 2)
    "brush2d_cyl": dict(
        text="""
lat : cyl : gradients : 2
lat : cyl : geometry : cylindrical
lat : cyl : n_layers_x : 14
lat : cyl : n_layers_y : 18
lat : cyl : lattice_type : simple_cubic
mon : X : freedom : pinned
mon : X : pinned_range : firstlayer_x
mon : A : freedom : free
mon : W : freedom : free
mon : A : chi_W : 0.5
mol : brush : composition : (X)1(A)29
mol : brush : freedom : restricted
mol : brush : n : 2.5
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(8, 0.05), (9, 0.5)],
        converged=[dict(method="DIIS", tol=1e-10, seed=10, amp=0.05)],
    ),
    # 2-gradient cylindrical brush on a rod, radial size a multiple of 8 (the
    # register-strip propagator path)
    "brush2d_rod_strip": dict(
        text="""
lat : cyl : gradients : 2
lat : cyl : geometry : cylindrical
lat : cyl : n_layers_x : 16
lat : cyl : n_layers_y : 24
lat : cyl : lattice_type : simple_cubic
mon : S : freedom : frozen
mon : S : frozen_range : 1,1;1,24
mon : X : freedom : pinned
mon : X : pinned_range : 2,1;2,24
mon : A : freedom : free
mon : W : freedom : free
mon : A : chi_W : 0.8
mon : S : chi_A : -0.3
mol : brush : composition : (X)1(A)39
mol : brush : freedom : restricted
mol : brush : n : 7.5398223686
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(21, 0.05), (22, 0.4)],
        converged=[dict(method="DIIS", tol=1e-10, seed=23, amp=0.05)],
    ),
    # 2-gradient cylindrical, hexagonal lattice, homopolymer near a wall
    "hex2d_cyl_strip": dict(
        text="""
lat : cyl : gradients : 2
lat : cyl : geometry : cylindrical
lat : cyl : n_layers_x : 8
lat : cyl : n_layers_y : 12
lat : cyl : lattice_type : hexagonal
mon : S : freedom : frozen
mon : S : frozen_range : 8,1;8,12
mon : A : freedom : free
mon : W : freedom : free
mon : S : chi_A : -0.8
mon : A : chi_W : 0.6
mol : pol : composition : (A)25
mol : pol : freedom : free
mol : pol : phibulk : 0.01
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(24, 0.3)],
    ),
    # 2-gradient planar, frozen particle, periodic y
    "planar2d_particle": dict(
        text="""
lat : L : gradients : 2
lat : L : n_layers_x : 12
lat : L : n_layers_y : 16
lat : L : lattice_type : simple_cubic
lat : L : lowerbound_y : periodic
lat : L : upperbound_y : periodic
mon : S : freedom : frozen
mon : S : frozen_range : 5,6;7,9
mon : A : freedom : free
mon : W : freedom : free
mon : S : chi_A : -1.0
mon : A : chi_W : 0.5
mol : pol : composition : (A)20
mol : pol : freedom : free
mol : pol : phibulk : 0.02
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(11, 0.3)],
    ),
    # 2-gradient planar, hexagonal 9-point stencil, frozen particle, periodic
    # y (radial size a multiple of 4: register-strip path)
    "hex2d_planar_particle": dict(
        text="""
lat : L : gradients : 2
lat : L : n_layers_x : 12
lat : L : n_layers_y : 16
lat : L : lattice_type : hexagonal
lat : L : lowerbound_y : periodic
lat : L : upperbound_y : periodic
mon : S : freedom : frozen
mon : S : frozen_range : 5,6;7,9
mon : A : freedom : free
mon : W : freedom : free
mon : S : chi_A : -1.0
mon : A : chi_W : 0.5
mol : pol : composition : (A)20
mol : pol : freedom : free
mol : pol : phibulk : 0.02
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(51, 0.3)],
        converged=[dict(method="DIIS", tol=1e-10, seed=52, amp=0.05)],
    ),
    # 2-gradient planar hexagonal brush grafted on a frozen wall (14 layers:
    # the site-strided path)
    "hex2d_planar_brush": dict(
        text="""
lat : L : gradients : 2
lat : L : n_layers_x : 14
lat : L : n_layers_y : 10
lat : L : lattice_type : hexagonal
mon : S : freedom : frozen
mon : S : frozen_range : 1,1;1,10
mon : X : freedom : pinned
mon : X : pinned_range : 2,1;2,10
mon : A : freedom : free
mon : W : freedom : free
mon : A : chi_W : 0.4
mon : S : chi_A : -0.2
mol : brush : composition : (X)1(A)24
mol : brush : freedom : restricted
mol : brush : n : 3
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(53, 0.3)],
    ),
    # 3-gradient periodic diblock box (small version of configuration 3)
    "diblock3d_periodic": dict(
        text="""
lat : box : gradients : 3
lat : box : n_layers_x : 8
lat : box : n_layers_y : 9
lat : box : n_layers_z : 10
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
mon : B : chi_W : 1.6
mon : A : chi_B : 1.6
mol : surf : composition : (A)10(B)10
mol : surf : freedom : restricted
mol : surf : theta : 36
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(12, 0.1), (13, 0.6)],
        converged=[dict(method="DIIS", tol=1e-9, seed=14, amp=0.1)],
    ),
    # 3-gradient mirror box with a frozen cube
    "particle3d_mirror": dict(
        text="""
lat : box : gradients : 3
lat : box : n_layers_x : 9
lat : box : n_layers_y : 8
lat : box : n_layers_z : 7
lat : box : lattice_type : simple_cubic
mon : S : freedom : frozen
mon : S : frozen_range : 1,1,1;3,3,3
mon : A : freedom : free
mon : W : freedom : free
mon : S : chi_A : -0.8
mon : A : chi_W : 0.4
mol : pol : composition : (A)15
mol : pol : freedom : free
mol : pol : phibulk : 0.03
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(15, 0.3)],
    ),
    # 3-gradient 27-point stencil (stencil_full false): periodic diblock box
    "diblock3d_27pt": dict(
        text="""
lat : box : gradients : 3
lat : box : n_layers_x : 8
lat : box : n_layers_y : 7
lat : box : n_layers_z : 9
lat : box : lattice_type : simple_cubic
lat : box : stencil_full : false
lat : box : lowerbound_x : periodic
lat : box : upperbound_x : periodic
lat : box : lowerbound_y : periodic
lat : box : upperbound_y : periodic
lat : box : lowerbound_z : periodic
lat : box : upperbound_z : periodic
mon : A : freedom : free
mon : B : freedom : free
mon : W : freedom : free
mon : B : chi_W : 1.6
mon : A : chi_B : 1.6
mol : surf : composition : (A)10(B)10
mol : surf : freedom : restricted
mol : surf : theta : 30
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(55, 0.2)],
        converged=[dict(method="DIIS", tol=1e-9, seed=56, amp=0.1)],
    ),
    # 3-gradient hexagonal lattice (the reference's unnormalised 64/40
    # stencil), mirror box with a frozen cube
    "hex3d_particle": dict(
        text="""
lat : box : gradients : 3
lat : box : n_layers_x : 7
lat : box : n_layers_y : 8
lat : box : n_layers_z : 6
lat : box : lattice_type : hexagonal
mon : S : freedom : frozen
mon : S : frozen_range : 1,1,1;2,3,2
mon : A : freedom : free
mon : W : freedom : free
mon : S : chi_A : -0.8
mon : A : chi_W : 0.4
mol : pol : composition : (A)12
mol : pol : freedom : free
mol : pol : phibulk : 0.03
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(57, 0.2)],
    ),
    # 3-gradient hexagonal lattice without stencil_full (4 : 2 : 1 / 56),
    # mirror in x, periodic in y and z
    "hex3d_27pt": dict(
        text="""
lat : box : gradients : 3
lat : box : n_layers_x : 6
lat : box : n_layers_y : 7
lat : box : n_layers_z : 8
lat : box : lattice_type : hexagonal
lat : box : stencil_full : false
lat : box : lowerbound_y : periodic
lat : box : upperbound_y : periodic
lat : box : lowerbound_z : periodic
lat : box : upperbound_z : periodic
mon : A : freedom : free
mon : W : freedom : free
mon : A : chi_W : 0.7
mol : pol : composition : (A)10
mol : pol : freedom : restricted
mol : pol : theta : 20
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(58, 0.3)],
    ),
    # refined 2-gradient cylindrical brush on a rod (FJC_choices 5,
    # hexagonal: the refined stencils need it)
    "fjc5_cyl2d_brush": dict(
        text="""
lat : cyl : gradients : 2
lat : cyl : geometry : cylindrical
lat : cyl : n_layers_x : 10
lat : cyl : n_layers_y : 8
lat : cyl : lattice_type : hexagonal
lat : cyl : FJC_choices : 5
mon : S : freedom : frozen
mon : S : frozen_range : 1,1;1,8
mon : X : freedom : pinned
mon : X : pinned_range : 2,1;2,8
mon : A : freedom : free
mon : W : freedom : free
mon : A : chi_W : 0.5
mon : S : chi_A : -0.3
mol : brush : composition : (X)1(A)20
mol : brush : freedom : restricted
mol : brush : n : 5
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(61, 0.2)],
        converged=[dict(method="DIIS", tol=1e-9, seed=62, amp=0.05)],
    ),
    # refined 2-gradient planar lattice (FJC_choices 5), frozen particle,
    # periodic y
    "fjc5_planar2d_particle": dict(
        text="""
lat : L : gradients : 2
lat : L : n_layers_x : 8
lat : L : n_layers_y : 9
lat : L : lattice_type : hexagonal
lat : L : FJC_choices : 5
lat : L : lowerbound_y : periodic
lat : L : upperbound_y : periodic
mon : S : freedom : frozen
mon : S : frozen_range : 3,3;4,5
mon : A : freedom : free
mon : W : freedom : free
mon : S : chi_A : -0.8
mon : A : chi_W : 0.5
mol : pol : composition : (A)15
mol : pol : freedom : free
mol : pol : phibulk : 0.02
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(63, 0.2)],
    ),
    # FJC_choices 7 on a 2-gradient cylindrical lattice (b/l 3)
    "fjc7_cyl2d": dict(
        text="""
lat : cyl : gradients : 2
lat : cyl : geometry : cylindrical
lat : cyl : n_layers_x : 6
lat : cyl : n_layers_y : 5
lat : cyl : offset_first_layer : 2
lat : cyl : lattice_type : hexagonal
lat : cyl : b/l : 3
mon : A : freedom : free
mon : W : freedom : free
mon : A : chi_W : 0.6
mol : pol : composition : (A)8
mol : pol : freedom : restricted
mol : pol : theta : 40
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(64, 0.2)],
    ),
    # FJC_choices 7 on a 2-gradient planar lattice, mirror boundaries
    "fjc7_planar2d": dict(
        text="""
lat : L : gradients : 2
lat : L : n_layers_x : 5
lat : L : n_layers_y : 6
lat : L : lattice_type : hexagonal
lat : L : FJC_choices : 7
mon : A : freedom : free
mon : W : freedom : free
mon : A : chi_W : 0.6
mol : pol : composition : (A)6
mol : pol : freedom : free
mol : pol : phibulk : 0.05
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(65, 0.2)],
    ),
    # refined 3-gradient hexagonal lattice (FJC_choices 5): a frozen wall
    # layer with an adsorbing polymer, periodic in y and z.  Evaluated at zero
    # fields and solved from zero fields (the reference misreads guess files
    # of refined 3-gradient lattices with several field blocks)
    "fjc5_wall3d": dict(
        text="""
lat : box : gradients : 3
lat : box : n_layers_x : 6
lat : box : n_layers_y : 3
lat : box : n_layers_z : 3
lat : box : lattice_type : hexagonal
lat : box : FJC_choices : 5
lat : box : lowerbound_y : periodic
lat : box : upperbound_y : periodic
lat : box : lowerbound_z : periodic
lat : box : upperbound_z : periodic
mon : S : freedom : frozen
mon : S : frozen_range : 1,1,1;1,3,3
mon : A : freedom : free
mon : W : freedom : free
mon : S : chi_A : -1.0
mon : A : chi_W : 0.5
mol : pol : composition : (A)10
mol : pol : freedom : free
mol : pol : phibulk : 0.02
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(66, 0.0)],
        converged=[dict(method="DIIS", tol=1e-9)],
    ),
    # refined 3-gradient hexagonal box with a frozen cube, mirror boundaries
    "fjc5_particle3d": dict(
        text="""
lat : box : gradients : 3
lat : box : n_layers_x : 5
lat : box : n_layers_y : 4
lat : box : n_layers_z : 4
lat : box : lattice_type : hexagonal
lat : box : FJC_choices : 5
mon : S : freedom : frozen
mon : S : frozen_range : 1,1,1;2,2,2
mon : A : freedom : free
mon : W : freedom : free
mon : S : chi_A : -0.6
mon : A : chi_W : 0.4
mol : pol : composition : (A)8
mol : pol : freedom : free
mol : pol : phibulk : 0.03
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(68, 0.0)],
    ),
    # second-order Markov on a 2-gradient cylindrical lattice (5 direction
    # components with the curved reflection weights): stiff brush on a rod
    "markov2_cyl2d_brush": dict(
        text="""
lat : cyl : gradients : 2
lat : cyl : geometry : cylindrical
lat : cyl : n_layers_x : 10
lat : cyl : n_layers_y : 12
lat : cyl : lattice_type : simple_cubic
mon : S : freedom : frozen
mon : S : frozen_range : 1,1;1,12
mon : X : freedom : pinned
mon : X : pinned_range : 2,1;2,12
mon : A : freedom : free
mon : W : freedom : free
mon : A : chi_W : 0.5
mon : S : chi_A : -0.3
mol : brush : composition : (X)1(A)20
mol : brush : freedom : restricted
mol : brush : n : 6
mol : brush : Markov : 2
mol : brush : k_stiff : 1
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(71, 0.2)],
        converged=[dict(method="DIIS", tol=1e-9, seed=72, amp=0.05)],
    ),
    # second-order Markov on a 2-gradient planar lattice next to a frozen
    # particle, periodic y, mixed with a first-order solvent (stencil_full
    # goes off for the whole lattice)
    "markov2_planar2d": dict(
        text="""
lat : L : gradients : 2
lat : L : n_layers_x : 9
lat : L : n_layers_y : 8
lat : L : lattice_type : simple_cubic
lat : L : lowerbound_y : periodic
lat : L : upperbound_y : periodic
mon : S : freedom : frozen
mon : S : frozen_range : 4,3;5,5
mon : A : freedom : free
mon : W : freedom : free
mon : S : chi_A : -0.9
mon : A : chi_W : 0.5
mol : pol : composition : (A)15
mol : pol : freedom : free
mol : pol : phibulk : 0.02
mol : pol : Markov : 2
mol : pol : k_stiff : 2
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(73, 0.2)],
        converged=[dict(method="DIIS", tol=1e-9, seed=74, amp=0.05)],
    ),
    # second-order Markov on a 3-gradient simple-cubic lattice (6 direction
    # components; stencil_full goes off for the whole lattice): polymer near
    # a frozen cube, mirror in x, periodic in y and z
    "markov2_box3d": dict(
        text="""
lat : box : gradients : 3
lat : box : n_layers_x : 7
lat : box : n_layers_y : 6
lat : box : n_layers_z : 5
lat : box : lattice_type : simple_cubic
lat : box : lowerbound_y : periodic
lat : box : upperbound_y : periodic
lat : box : lowerbound_z : periodic
lat : box : upperbound_z : periodic
mon : S : freedom : frozen
mon : S : frozen_range : 1,1,1;2,3,5
mon : A : freedom : free
mon : W : freedom : free
mon : S : chi_A : -0.8
mon : A : chi_W : 0.5
mol : pol : composition : (A)12
mol : pol : freedom : free
mol : pol : phibulk : 0.03
mol : pol : Markov : 2
mol : pol : k_stiff : 1.5
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(75, 0.2)],
        converged=[dict(method="DIIS", tol=1e-9, seed=76, amp=0.05)],
    ),
}
