"""Ring molecules (`mol : X : ring : true`), for reference fixtures and
parity tests.  The first and last blocks of a linear chain, one segment each
of the same type, are merged: a chain of the composition's length less one
that starts and ends on the same site.  Cases: rings adsorbing on a planar
wall next to free linear chains, rings around a spherical particle, rings
in a 2-gradient cylinder and in a periodic 3-gradient box.
"""

RING = {
    "ring1d_planar": dict(
        text="""
lat : flat : gradients : 1
lat : flat : n_layers : 40
lat : flat : lattice_type : simple_cubic
lat : flat : lowerbound : surface
mon : S : freedom : frozen
mon : S : frozen_range : lowerbound
mon : A : freedom : free
mon : B : freedom : free
mon : W : freedom : free
mon : S : chi_A : -1.0
mon : B : chi_W : 0.4
mol : ring : composition : (A)1(B)8(A)1(B)9(A)1
mol : ring : ring : true
mol : ring : freedom : restricted
mol : ring : theta : 3
mol : lin : composition : (B)12
mol : lin : freedom : free
mol : lin : phibulk : 0.01
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(401, 0.3)],
        converged=[dict(method="DIIS", tol=1e-8)],
    ),
    "ring1d_sphere": dict(
        text="""
lat : ball : gradients : 1
lat : ball : geometry : spherical
lat : ball : n_layers : 30
lat : ball : offset_first_layer : 3
lat : ball : lattice_type : simple_cubic
lat : ball : lowerbound : surface
mon : S : freedom : frozen
mon : S : frozen_range : lowerbound
mon : A : freedom : free
mon : W : freedom : free
mon : S : chi_A : -1.5
mon : A : chi_W : 0.5
mol : ring : composition : (A)1(A)14(A)1
mol : ring : ring : true
mol : ring : freedom : free
mol : ring : phibulk : 0.002
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(402, 0.3)],
        converged=[dict(method="DIIS", tol=1e-8)],
    ),
    "ring2d_cyl": dict(
        text="""
lat : cyl : gradients : 2
lat : cyl : geometry : cylindrical
lat : cyl : n_layers_x : 10
lat : cyl : n_layers_y : 12
lat : cyl : lattice_type : simple_cubic
mon : A : freedom : free
mon : W : freedom : free
mon : A : chi_W : 0.5
mol : ring : composition : (A)1(A)10(A)1
mol : ring : ring : true
mol : ring : freedom : restricted
mol : ring : theta : 6
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(403, 0.3)],
        converged=[dict(method="DIIS", tol=1e-8)],
    ),
    "ring3d_box": dict(
        text="""
lat : box : gradients : 3
lat : box : n_layers_x : 6
lat : box : n_layers_y : 5
lat : box : n_layers_z : 7
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
mon : A : chi_W : 0.8
mon : B : chi_W : 0.3
mol : ring : composition : (A)1(B)4(A)2(B)3(A)1
mol : ring : ring : true
mol : ring : freedom : restricted
mol : ring : theta : 2
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(404, 0.3)],
        converged=[dict(method="DIIS", tol=1e-8)],
    ),
}
