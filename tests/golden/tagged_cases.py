"""Tagged segments, for reference fixtures and parity tests.  A tagged
segment sits on the sites of its tagged_range and nowhere else: its
statistical weight is the mask (no field), its density is the mask, and the
molecule that holds it is normalised on the number of tagged sites, with
the tagged segment left out of the chain length in F, Omega and mu.  Cases:
an end-tagged brush next to a wall (one chain per site of the planar
layer), a chain tagged in its middle on one shell of a spherical lattice in
a solution of free chains of the same monomers, and a chain tagged at one
site of a 2-gradient cylinder.
"""

TAGGED = {
    "tag1d_brush": dict(
        text="""
lat : flat : gradients : 1
lat : flat : n_layers : 25
lat : flat : lattice_type : simple_cubic
lat : flat : lowerbound : surface
mon : S : freedom : frozen
mon : S : frozen_range : lowerbound
mon : X : freedom : tagged
mon : X : tagged_range : 1;1
mon : A : freedom : free
mon : W : freedom : free
mon : S : chi_A : -0.3
mon : A : chi_W : 0.5
mon : X : chi_W : 0.3
mol : brush : composition : (X)1(A)6
mol : brush : freedom : tagged
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(301, 0.3)],
        converged=[dict(method="DIIS", tol=1e-8)],
    ),
    "tag1d_middle": dict(
        text="""
lat : ball : gradients : 1
lat : ball : geometry : spherical
lat : ball : n_layers : 30
lat : ball : lattice_type : simple_cubic
mon : X : freedom : tagged
mon : X : tagged_range : 8;8
mon : A : freedom : free
mon : B : freedom : free
mon : W : freedom : free
mon : A : chi_W : 0.8
mon : B : chi_W : 0.2
mon : X : chi_A : -0.4
mol : tag : composition : (A)12(X)1(B)12
mol : tag : freedom : tagged
mol : pol : composition : (A)20
mol : pol : freedom : free
mol : pol : phibulk : 0.01
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(302, 0.3)],
        converged=[dict(method="DIIS", tol=1e-8)],
    ),
    "tag2d_cyl": dict(
        text="""
lat : cyl : gradients : 2
lat : cyl : geometry : cylindrical
lat : cyl : n_layers_x : 12
lat : cyl : n_layers_y : 16
lat : cyl : lattice_type : simple_cubic
mon : X : freedom : tagged
mon : X : tagged_range : 1,8;1,8
mon : A : freedom : free
mon : W : freedom : free
mon : A : chi_W : 0.7
mol : tag : composition : (X)1(A)24
mol : tag : freedom : tagged
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(303, 0.3)],
        converged=[dict(method="DIIS", tol=1e-8)],
    ),
}
