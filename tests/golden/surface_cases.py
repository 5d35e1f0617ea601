"""Surface boundaries on 2- and 3-gradient lattices, for reference fixtures
and parity tests.  A surface ghost keeps its own value (no copy from the
interior), so what the reference computes on it over its flattened arrays
reaches the interior; the library runs these systems with the chains over
every site.  Cases: a 2-gradient planar wall (periodic along the wall) and
a 2-gradient cylinder closed by surfaces at both ends.  (The reference
offers no surface boundary on 3-gradient lattices.)
"""

SURFACE = {
    "surf2d_planar": dict(
        text="""
lat : L : gradients : 2
lat : L : n_layers_x : 10
lat : L : n_layers_y : 12
lat : L : lattice_type : simple_cubic
lat : L : lowerbound_x : surface
lat : L : lowerbound_y : periodic
lat : L : upperbound_y : periodic
mon : S : freedom : frozen
mon : S : frozen_range : lowerbound_x
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
        eval=[(201, 0.3)],
        converged=[dict(method="DIIS", tol=1e-8)],
    ),
    "surf2d_cyl": dict(
        text="""
lat : cyl : gradients : 2
lat : cyl : geometry : cylindrical
lat : cyl : n_layers_x : 12
lat : cyl : n_layers_y : 10
lat : cyl : lattice_type : simple_cubic
lat : cyl : lowerbound_y : surface
lat : cyl : upperbound_y : surface
mon : S : freedom : frozen
mon : S : frozen_range : lowerbound_y
mon : T : freedom : frozen
mon : T : frozen_range : upperbound_y
mon : A : freedom : free
mon : B : freedom : free
mon : W : freedom : free
mon : S : chi_A : -0.8
mon : T : chi_B : -0.8
mon : A : chi_B : 0.6
mol : pol : composition : (A)8(B)8
mol : pol : freedom : restricted
mol : pol : theta : 30
mol : water : composition : (W)1
mol : water : freedom : solvent
""",
        eval=[(202, 0.3)],
        converged=[dict(method="DIIS", tol=1e-8)],
    ),
}
