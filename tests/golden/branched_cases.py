"""Branched molecules (square brackets in the composition) for reference
fixtures and parity tests.

Each case is `.in` text in the reference's input format; `eval` lists
(seed, amplitude) pairs of uniform random fields at which the residual is
evaluated, `converged` the DIIS runs (from zero fields, or from seeded
fields where zero fields are already the homogeneous solution).  The compositions
cover: side groups on a branch point inside the backbone, adjacent groups
on one branch point, a group nested in a side chain, a chain that starts
with a branch point (the reference then sums GN over the row in front of
the side group), free and restricted molecules, and 1-, 2- and 3-gradient
lattices (planar and cylindrical); symmetric and asymmetric
dendrimers (@dend): stars and two-generation dendrimers, with multi-block
spacers; and combs (@comb) with one- and two-block teeth.
"""

WALL_1D = """
lat : flat : gradients : 1
lat : flat : n_layers : 40
lat : flat : lattice_type : simple_cubic
lat : flat : lowerbound : surface
mon : S : freedom : frozen
mon : S : frozen_range : lowerbound
"""

BRANCHED = {
    # a CTA-like surfactant: a tail, a head segment with two side segments,
    # and the junction as the chain's last segment
    "branch1d_head": dict(
        text=WALL_1D + """
mon : C : freedom : free
mon : N : freedom : free
mon : W : freedom : free
mon : S : chi_C : -0.5
mon : C : chi_W : 1.2
mon : N : chi_W : -0.3
mol : surf : composition : (C)8(N)1[(N)1][(N)1](N)1
mol : surf : freedom : restricted
mol : surf : theta : 3.0
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(71, 0.3)],
        converged=[dict(method="DIIS", tol=1e-7)],
    ),
    # a comb: two side chains on junctions inside the backbone
    "branch1d_comb": dict(
        text=WALL_1D + """
mon : A : freedom : free
mon : B : freedom : free
mon : W : freedom : free
mon : S : chi_A : -1.0
mon : B : chi_W : 0.6
mol : comb : composition : (A)4[(B)3](A)1(A)4[(B)3](A)1(A)3
mol : comb : freedom : free
mol : comb : phibulk : 0.002
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(72, 0.3)],
        converged=[dict(method="DIIS", tol=1e-7)],
    ),
    # a chain that starts with a one-segment branch point: the reference's
    # propagation then sums GN over the row in front of the side group
    "branch1d_start": dict(
        text=WALL_1D + """
mon : C : freedom : free
mon : N : freedom : free
mon : W : freedom : free
mon : S : chi_C : -0.5
mon : C : chi_W : 1.2
mon : N : chi_W : -0.3
mol : surf : composition : (N)1[(N)2](C)8
mol : surf : freedom : restricted
mol : surf : theta : 3.0
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(77, 0.3)],
        converged=[dict(method="DIIS", tol=1e-7)],
    ),
    # a side chain that carries a group of its own
    "branch1d_nested": dict(
        text=WALL_1D + """
mon : A : freedom : free
mon : B : freedom : free
mon : C : freedom : free
mon : W : freedom : free
mon : S : chi_A : -0.8
mon : C : chi_W : 0.8
mol : tree : composition : (A)3[(B)2[(C)2](B)1](A)1(A)2
mol : tree : freedom : restricted
mol : tree : theta : 2.0
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(73, 0.3)],
        converged=[dict(method="DIIS", tol=1e-7)],
    ),
    # a branched surfactant in a cylindrical micelle (curved 1-gradient)
    "branch1d_cyl": dict(
        text="""
lat : flat : gradients : 1
lat : flat : geometry : cylindrical
lat : flat : n_layers : 30
lat : flat : lattice_type : simple_cubic
mon : C : freedom : free
mon : N : freedom : free
mon : W : freedom : free
mon : C : chi_W : 1.6
mon : N : chi_W : -0.2
mol : surf : composition : (C)6[(N)1](N)1(N)1[(N)1][(N)1](N)1
mol : surf : freedom : restricted
mol : surf : theta : 60
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(74, 0.2)],
        converged=[dict(method="DIIS", tol=1e-7, seed=84, amp=0.2)],
    ),
    # a lipid-like molecule in a 2-gradient planar box
    "branch2d_lipid": dict(
        text="""
lat : flat : gradients : 2
lat : flat : n_layers_x : 10
lat : flat : n_layers_y : 12
lat : flat : lattice_type : simple_cubic
lat : flat : lowerbound_x : mirror
lat : flat : upperbound_x : mirror
lat : flat : lowerbound_y : periodic
lat : flat : upperbound_y : periodic
mon : C : freedom : free
mon : O : freedom : free
mon : N : freedom : free
mon : W : freedom : free
mon : C : chi_W : 1.4
mon : O : chi_W : 0.2
mon : N : chi_W : -0.4
mol : lip : composition : (C)6[(O)1](O)1(C)2[(N)1][(N)1](N)1
mol : lip : freedom : restricted
mol : lip : theta : 40
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(75, 0.2)],
        converged=[dict(method="DIIS", tol=1e-7, seed=85, amp=0.2)],
    ),
    # a 3-gradient periodic box
    "branch3d_box": dict(
        text="""
lat : box : gradients : 3
lat : box : n_layers_x : 8
lat : box : n_layers_y : 8
lat : box : n_layers_z : 8
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
mon : A : chi_W : 1.2
mon : B : chi_W : 0.3
mol : br : composition : (A)3[(B)2](A)1(A)3
mol : br : freedom : restricted
mol : br : theta : 60
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(76, 0.2)],
        converged=[dict(method="DIIS", tol=1e-7, seed=86, amp=0.2)],
    ),
    # symmetric dendrimers (@dend): a four-arm star at a wall, a
    # two-generation dendrimer in a cylindrical micelle-like confinement,
    # and both in 2- and 3-gradient boxes
    "dend1d_star": dict(
        text=WALL_1D + """
mon : A : freedom : free
mon : B : freedom : free
mon : W : freedom : free
mon : S : chi_B : -0.6
mon : B : chi_W : 0.5
mol : star : composition : @dend(A,(B)8,4)
mol : star : freedom : restricted
mol : star : theta : 4.0
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(91, 0.3)],
        converged=[dict(method="DIIS", tol=1e-7)],
    ),
    "dend1d_gen2": dict(
        text=WALL_1D + """
mon : A : freedom : free
mon : B : freedom : free
mon : C : freedom : free
mon : W : freedom : free
mon : S : chi_A : -0.4
mon : C : chi_W : 0.9
mol : dend : composition : @dend(A,(B)3,3;A,(B)2(C)1,2)
mol : dend : freedom : free
mol : dend : phibulk : 0.003
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(92, 0.3)],
        converged=[dict(method="DIIS", tol=1e-7)],
    ),
    "dend2d_star": dict(
        text="""
lat : flat : gradients : 2
lat : flat : n_layers_x : 10
lat : flat : n_layers_y : 12
lat : flat : lattice_type : simple_cubic
lat : flat : lowerbound_x : mirror
lat : flat : upperbound_x : mirror
lat : flat : lowerbound_y : periodic
lat : flat : upperbound_y : periodic
mon : A : freedom : free
mon : B : freedom : free
mon : W : freedom : free
mon : B : chi_W : 0.8
mol : star : composition : @dend(A,(B)4,3)
mol : star : freedom : restricted
mol : star : theta : 30
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(93, 0.2)],
        converged=[dict(method="DIIS", tol=1e-7, seed=94, amp=0.1)],
    ),
    "dend3d_gen2": dict(
        text="""
lat : box : gradients : 3
lat : box : n_layers_x : 8
lat : box : n_layers_y : 8
lat : box : n_layers_z : 8
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
mon : B : chi_W : 1.1
mol : dend : composition : @dend(A,(B)2,3;A,(B)2,2)
mol : dend : freedom : restricted
mol : dend : theta : 60
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(95, 0.2)],
        converged=[dict(method="DIIS", tol=1e-7, seed=96, amp=0.2)],
    ),
    # combs (@comb): head, repeats of branch point + tooth + spacer, tail
    "comb1d_wall": dict(
        text=WALL_1D + """
mon : A : freedom : free
mon : B : freedom : free
mon : W : freedom : free
mon : S : chi_A : -0.7
mon : B : chi_W : 0.6
mol : comb : composition : @comb((A)3;A,(B)3,(A)2,3;(A)4)
mol : comb : freedom : restricted
mol : comb : theta : 3.0
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(97, 0.3)],
        converged=[dict(method="DIIS", tol=1e-7)],
    ),
    "comb1d_free": dict(
        text=WALL_1D + """
mon : A : freedom : free
mon : B : freedom : free
mon : C : freedom : free
mon : W : freedom : free
mon : S : chi_B : -0.5
mon : C : chi_W : 0.8
mol : comb : composition : @comb((A)1;A,(B)2(C)1,(A)3,2;(A)2)
mol : comb : freedom : free
mol : comb : phibulk : 0.004
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(98, 0.3)],
        converged=[dict(method="DIIS", tol=1e-7)],
    ),
    "comb3d_box": dict(
        text="""
lat : box : gradients : 3
lat : box : n_layers_x : 8
lat : box : n_layers_y : 8
lat : box : n_layers_z : 8
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
mon : B : chi_W : 1.0
mol : comb : composition : @comb((A)2;A,(B)2,(A)2,2;(A)2)
mol : comb : freedom : restricted
mol : comb : theta : 50
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(99, 0.2)],
        converged=[dict(method="DIIS", tol=1e-7, seed=100, amp=0.1)],
    ),
    # asymmetric dendrimers: arm types of different length and count per
    # generation (a star with one different arm; two generations)
    "adend1d_star": dict(
        text=WALL_1D + """
mon : A : freedom : free
mon : B : freedom : free
mon : C : freedom : free
mon : W : freedom : free
mon : S : chi_C : -1.0
mon : B : chi_W : 0.4
mol : star : composition : @dend(A,(B)6,3,(B)5(C)1,1)
mol : star : freedom : restricted
mol : star : theta : 3.0
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(101, 0.3)],
        converged=[dict(method="DIIS", tol=1e-7)],
    ),
    "adend1d_gen2": dict(
        text=WALL_1D + """
mon : A : freedom : free
mon : B : freedom : free
mon : C : freedom : free
mon : W : freedom : free
mon : S : chi_A : -0.5
mon : C : chi_W : 0.7
mol : dend : composition : @dend(A,(B)2,2,(C)3,1;A,(B)2,2)
mol : dend : freedom : free
mol : dend : phibulk : 0.002
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(102, 0.3)],
        converged=[dict(method="DIIS", tol=1e-7)],
    ),
    "adend3d_box": dict(
        text="""
lat : box : gradients : 3
lat : box : n_layers_x : 8
lat : box : n_layers_y : 8
lat : box : n_layers_z : 8
lat : box : lattice_type : simple_cubic
lat : box : lowerbound_x : periodic
lat : box : upperbound_x : periodic
lat : box : lowerbound_y : periodic
lat : box : upperbound_y : periodic
lat : box : lowerbound_z : periodic
lat : box : upperbound_z : periodic
mon : A : freedom : free
mon : B : freedom : free
mon : C : freedom : free
mon : W : freedom : free
mon : C : chi_W : 1.0
mol : dend : composition : @dend(A,(B)2,2,(C)2,2;A,(B)1,2)
mol : dend : freedom : restricted
mol : dend : theta : 50
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(103, 0.2)],
        converged=[dict(method="DIIS", tol=1e-7, seed=104, amp=0.1)],
    ),
}
