"""Invariants that pin the oracle independently of the reference.

Closed forms of the lattice operators (eigenmodes, row sums, detailed
balance), composition-law identities, the chemical-potential closed form and
the DIIS least-squares step.  Each would catch a dropped term, a wrong
weight, a wrong boundary remap or a transposed operand.
"""
import numpy as np
import pytest

from tests.oracle_lib import Oracle, solve_ones

LAT1 = """
lat : L : gradients : 1
lat : L : n_layers : {n}
lat : L : lattice_type : {lt}
lat : L : geometry : {geo}
{extra}
mon : A : freedom : free
mon : W : freedom : free
mon : A : chi_W : 0.5
mol : pol : composition : (A)10
mol : pol : freedom : free
mol : pol : phibulk : 0.01
mol : water : composition : (W)1
mol : water : freedom : solvent
"""


def lat1(n=24, lt="simple_cubic", geo="planar", extra=""):
    return Oracle(LAT1.format(n=n, lt=lt, geo=geo, extra=extra))


@pytest.mark.parametrize("lt,lam", [("simple_cubic", 1 / 6), ("hexagonal", 1 / 4)])
def test_planar_mirror_eigenmodes(lt, lam):
    """cos(k(x-1/2)), k = pi j / MX, are eigenvectors of the mirror-bounded
    (lambda, 1-2 lambda, lambda) stencil with eigenvalue 1-2lambda+2lambda cos k."""
    o = lat1(n=20, lt=lt)
    MX = 20
    xs = np.arange(1, MX + 1)
    for j in (0, 1, 5, 19):
        k = np.pi * j / MX
        v = np.zeros(o.M)
        v[1:MX + 1] = np.cos(k * (xs - 0.5))
        w = o.apply(v)
        ev = 1 - 2 * lam + 2 * lam * np.cos(k)
        np.testing.assert_allclose(w[1:MX + 1], ev * v[1:MX + 1], atol=1e-13)
        # Side uses the same weights
        np.testing.assert_allclose(o.apply(v, side=True)[1:MX + 1], ev * v[1:MX + 1], atol=1e-13)


def test_planar_periodic_eigenmodes():
    o = lat1(n=16, extra="lat : L : lowerbound : periodic\nlat : L : upperbound : periodic")
    MX, lam = 16, 1 / 6
    xs = np.arange(1, MX + 1)
    for j in (1, 3, 8):
        k = 2 * np.pi * j / MX
        v = np.zeros(o.M)
        v[1:MX + 1] = np.sin(k * xs + 0.3)
        ev = 1 - 2 * lam + 2 * lam * np.cos(k)
        np.testing.assert_allclose(o.apply(v)[1:MX + 1], ev * v[1:MX + 1], atol=1e-13)


@pytest.mark.parametrize("geo,extra", [("cylindrical", ""), ("cylindrical", "lat : L : offset_first_layer : 4"),
                                       ("spherical", ""), ("spherical", "lat : L : offset_first_layer : 2.5")])
def test_curved_rows_sum_to_one_and_detailed_balance(geo, extra):
    """Curved-lattice weights: W 1 = 1 (mirror bounds) and W is symmetric in
    the L-weighted inner product (L(x) lambda+(x) = L(x+1) lambda-(x+1))."""
    o = lat1(n=18, geo=geo, extra=extra)
    MX = 18
    one = np.ones(o.M)
    np.testing.assert_allclose(o.apply(one)[1:MX + 1], 1.0, atol=1e-14)
    rng = np.random.default_rng(0)
    u = np.zeros(o.M)
    v = np.zeros(o.M)
    u[1:MX + 1] = rng.standard_normal(MX)
    v[1:MX + 1] = rng.standard_normal(MX)
    # site volumes from the reference formulas with its pi constant
    PIE = 3.14159265
    off = 0.0 if not extra else float(extra.split(":")[-1])
    r = off + np.arange(1, MX + 1)
    L = PIE * (r ** 2 - (r - 1) ** 2) if geo == "cylindrical" else 4 / 3 * PIE * (r ** 3 - (r - 1) ** 3)
    Wu = o.apply(u)[1:MX + 1]
    Wv = o.apply(v)[1:MX + 1]
    lhs = np.sum(L * u[1:MX + 1] * Wv)
    rhs = np.sum(L * v[1:MX + 1] * Wu)
    assert abs(lhs - rhs) < 1e-11 * np.sum(L)


BOX = """
lat : box : gradients : 3
lat : box : n_layers_x : 6
lat : box : n_layers_y : 5
lat : box : n_layers_z : 4
lat : box : lattice_type : simple_cubic
lat : box : lowerbound_x : periodic
lat : box : upperbound_x : periodic
lat : box : lowerbound_y : {by}
lat : box : upperbound_y : {by}
lat : box : lowerbound_z : periodic
lat : box : upperbound_z : periodic
mon : A : freedom : free
mon : W : freedom : free
mol : pol : composition : (A)5
mol : pol : freedom : free
mol : pol : phibulk : 0.1
mol : water : composition : (W)1
mol : water : freedom : solvent
"""


@pytest.mark.parametrize("by", ["periodic", "mirror"])
def test_3d_separable_eigenmodes(by):
    """Products of 1-D modes are eigenvectors of the 6-point stencil with
    eigenvalue (cos kx + cos ky + cos kz)/3."""
    o = Oracle(BOX.format(by=by))
    MX, MY, MZ = 6, 5, 4
    x = np.arange(1, MX + 1)[:, None, None]
    y = np.arange(1, MY + 1)[None, :, None]
    z = np.arange(1, MZ + 1)[None, None, :]
    kx, kz = 2 * np.pi * 2 / MX, 2 * np.pi * 1 / MZ
    if by == "periodic":
        ky = 2 * np.pi * 1 / MY
        fy = np.cos(ky * y)
    else:
        ky = np.pi * 2 / MY
        fy = np.cos(ky * (y - 0.5))
    mode = np.cos(kx * x) * fy * np.sin(kz * z + 0.1)
    v = np.zeros((MX + 2, MY + 2, MZ + 2))
    v[1:-1, 1:-1, 1:-1] = mode
    w = o.apply(v.ravel()).reshape(MX + 2, MY + 2, MZ + 2)[1:-1, 1:-1, 1:-1]
    ev = (np.cos(kx) + np.cos(ky) + np.cos(kz)) / 3
    np.testing.assert_allclose(w, ev * mode, atol=1e-13)


PLANAR2D = """
lat : L : gradients : 2
lat : L : n_layers_x : 7
lat : L : n_layers_y : 9
lat : L : lattice_type : simple_cubic
lat : L : lowerbound_y : periodic
lat : L : upperbound_y : periodic
mon : A : freedom : free
mon : W : freedom : free
mol : pol : composition : (A)5
mol : pol : freedom : free
mol : pol : phibulk : 0.1
mol : water : composition : (W)1
mol : water : freedom : solvent
"""


def test_2d_planar_separable_weights():
    """The 9-point stencil is (1,4,1)/6 (x) (1,4,1)/6: eigenvalue
    (4+2cos kx)(4+2cos ky)/36 for mirror-x / periodic-y modes."""
    o = Oracle(PLANAR2D)
    MX, MY = 7, 9
    kx, ky = np.pi * 3 / MX, 2 * np.pi * 2 / MY
    x = np.arange(1, MX + 1)[:, None]
    y = np.arange(1, MY + 1)[None, :]
    mode = np.cos(kx * (x - 0.5)) * np.cos(ky * y + 0.2)
    v = np.zeros((MX + 2, MY + 2))
    v[1:-1, 1:-1] = mode
    w = o.apply(v.ravel()).reshape(MX + 2, MY + 2)[1:-1, 1:-1]
    np.testing.assert_allclose(w, (4 + 2 * np.cos(kx)) * (4 + 2 * np.cos(ky)) / 36 * mode, atol=1e-13)


CYL2D = """
lat : cyl : gradients : 2
lat : cyl : geometry : cylindrical
lat : cyl : n_layers_x : 8
lat : cyl : n_layers_y : 6
lat : cyl : lattice_type : simple_cubic
mon : A : freedom : free
mon : W : freedom : free
mol : pol : composition : (A)5
mol : pol : freedom : free
mol : pol : phibulk : 0.1
mol : water : composition : (W)1
mol : water : freedom : solvent
"""


def test_2d_cylindrical_is_radial_times_axial():
    """The cylindrical 2-D stencil factorises into the 1-D cylindrical radial
    operator (with lambda scaled so the radial row is (lambda-, 4/6, lambda+))
    times the axial (1,4,1)/6: a product f(x) g(y) maps to (W_r f)(W_z g)."""
    o2 = Oracle(CYL2D)
    MX, MY = 8, 6
    rng = np.random.default_rng(3)
    f = rng.standard_normal(MX)
    ky = np.pi * 1 / MY
    gy = np.cos(ky * (np.arange(1, MY + 1) - 0.5))
    v = np.zeros((MX + 2, MY + 2))
    v[1:-1, 1:-1] = f[:, None] * gy[None, :]
    w = o2.apply(v.ravel()).reshape(MX + 2, MY + 2)[1:-1, 1:-1]
    PIE = 3.14159265
    r = np.arange(1, MX + 1).astype(float)
    L = PIE * (r ** 2 - (r - 1) ** 2)
    lp = 2 * PIE * r / L / 6
    lm = 2 * PIE * (r - 1) / L / 6
    fb = np.concatenate([[f[0]], f, [f[-1]]])  # mirror bounds
    Wf = lm * fb[:-2] + (4 / 6) * fb[1:-1] + lp * fb[2:]
    np.testing.assert_allclose(w, Wf[:, None] * ((4 + 2 * np.cos(ky)) / 6 * gy)[None, :], atol=1e-13)


def test_markov2_straight_and_turn_weights():
    """Markov-2 step on a uniform field: each direction component keeps the
    total weight P0 + 4 P1 = 1 (normalised bending weights, no backfolding),
    and with k_stiff = 0 P0 = P1 = 1/5."""
    text = LAT1.format(n=12, lt="simple_cubic", geo="planar", extra="").replace(
        "mol : pol : phibulk : 0.01", "mol : pol : phibulk : 0.01\nmol : pol : Markov : 2\nmol : pol : k_stiff : 0")
    o = Oracle(text)
    X = np.ones((3, o.M))
    out = o.apply_markov2(X)
    np.testing.assert_allclose(out[:, 1:13], 1.0, atol=1e-14)
    # a walker in +x only: next +x weight P0 on layer x+1, lateral 4P1 there... P0 = P1 = 1/5
    X = np.zeros((3, o.M))
    X[0, 6] = 1.0
    out = o.apply_markov2(X)
    assert abs(out[0, 7] - 0.2) < 1e-15      # straight on
    assert abs(out[1, 6] - 0.2) < 1e-15      # turn: lateral component at the same layer (one of 4 lateral)
    assert abs(out[2, 5]) < 1e-15            # no backfolding


def test_composition_identity_and_mu_closed_form(golden):
    """For a homopolymer, sum_x Gf(x,s) Gb(x,s) / G1(x) = GN for every s, so
    sum_x phi_pol(x) = n N = theta.  At zero field with no wall the chemical
    potential is the Flory-Huggins value."""
    fx = golden["cfg1_planar_adsorption"]
    o = Oracle(fx["text"])
    x = np.array(fx["evaluation"][0]["x"])
    o.residual(x)
    r = o.results()
    theta = r["mol"]["pol"]["theta"]
    assert abs(np.sum(o.interior(r["mol_phi"]["pol"])) - theta) < 1e-10 * theta
    assert abs(theta - r["mol"]["pol"]["n"] * 100) < 1e-10 * theta
    # Flory-Huggins mu of the bulk: ln(phi) + 1 - N(phi/N + phiW) + N chi phiW^2
    phi, N, chi = 0.01, 100, 0.5
    mu = np.log(phi) + 1 - N * (phi / N + (1 - phi)) + N * chi * (1 - phi) ** 2
    assert abs(r["mol"]["pol"]["Mu"] - mu) < 1e-12 * abs(mu)


def test_homogeneous_bulk_is_a_fixed_point():
    """Without walls the homogeneous bulk (x = 0) has g = 0 and Omega = 0."""
    o = lat1(n=30)
    g, res = o.residual(np.zeros(o.nvar))
    assert res < 1e-13
    assert abs(o.results()["grand_potential"]) < 1e-12


def test_diis_solve_matches_numpy():
    """DIIS coefficients c = A^-1 1 / sum(A^-1 1) (the reference returns 1
    directly for a 1 x 1 system)."""
    rng = np.random.default_rng(7)
    for n in (1, 2, 3, 6, 10):
        B = rng.standard_normal((n + 3, n))
        A = B.T @ B + 1e-3 * np.eye(n)
        c = solve_ones(A)
        ref = np.linalg.solve(A, np.ones(n))
        np.testing.assert_allclose(c / c.sum(), ref / ref.sum(), rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("fjc", [2, 3, 4])
def test_refined_periodic_eigenmodes(fjc):
    """Refined planar lattice (FJC = 2 fjc + 1): the stencil has weight
    1/(2(FJC-1)) at distance fjc and 1/(FJC-1) at distances 0..fjc-1, so
    sin(kx + c) with k = 2 pi j / MX (MX = fjc n) is an eigenvector of the
    periodic operator with eigenvalue w0 + 2 sum_d w_d cos(k d); the same
    weights serve the site average."""
    n = 12
    o = lat1(n=n, lt="hexagonal",
             extra="lat : L : FJC_choices : %d\nlat : L : lowerbound : periodic\nlat : L : upperbound : periodic"
             % (2 * fjc + 1))
    F, MX = 2 * fjc + 1, fjc * n
    assert o.M == MX + 2 * fjc
    w = [1.0 / (F - 1)] * fjc + [0.5 / (F - 1)]          # distances 0..fjc
    xs = np.arange(fjc, MX + fjc)
    for j in (1, 4, n):
        k = 2 * np.pi * j / MX
        v = np.zeros(o.M)
        v[fjc:MX + fjc] = np.sin(k * xs + 0.3)
        ev = w[0] + 2 * sum(w[d] * np.cos(k * d) for d in range(1, fjc + 1))
        np.testing.assert_allclose(o.apply(v)[fjc:MX + fjc], ev * v[fjc:MX + fjc], atol=1e-13)
        np.testing.assert_allclose(o.apply(v, side=True)[fjc:MX + fjc], ev * v[fjc:MX + fjc], atol=1e-13)


@pytest.mark.parametrize("geo,extra", [("planar", ""), ("cylindrical", ""), ("spherical", ""),
                                       ("cylindrical", "lat : L : offset_first_layer : 3")])
def test_refined_rows_sum_to_one_and_zero_field_gn(geo, extra):
    """Refined lattices (FJC 5): with mirror bounds the operator keeps a
    constant (every row of weights sums to one, the self weight absorbing the
    remainder), and on the planar lattice a chain in zero field has
    GN = volume = n_layers (sums carry the 1/fjc site weight)."""
    n = 15
    o = lat1(n=n, lt="hexagonal", geo=geo, extra="lat : L : FJC_choices : 5\n" + extra)
    f = 2
    one = np.ones(o.M)
    np.testing.assert_allclose(o.apply(one)[f:f * n + f], 1.0, atol=1e-14)
    if geo == "planar":
        o.residual(np.zeros(o.nvar))
        assert abs(o.results()["mol"]["pol"]["GN"] - n) < 1e-12 * n


@pytest.mark.parametrize("fjc,kst", [(1, 0.0), (2, 0.0), (2, 1.5), (3, 0.7)])
def test_hexagonal_markov2_keeps_a_constant(fjc, kst):
    """Planar hexagonal Markov 2 with FJC components: the bending weights are
    normalised so that every component's incoming weights sum to one, hence
    in zero field with mirror bounds all components stay 1 and
    GN = (G_0 + 2 sum G_inner + G_last) / (4 fjc) equals the volume."""
    n = 14
    text = LAT1.format(n=n, lt="hexagonal", geo="planar",
                       extra="lat : L : FJC_choices : %d" % (2 * fjc + 1)).replace(
        "mol : pol : phibulk : 0.01", "mol : pol : phibulk : 0.01\nmol : pol : Markov : 2\nmol : pol : k_stiff : %g" % kst)
    o = Oracle(text)
    o.residual(np.zeros(o.nvar))
    assert abs(o.results()["mol"]["pol"]["GN"] - n) < 1e-11 * n


def test_surface_needs_a_frozen_layer():
    """A surface boundary without a frozen segment on it is refused, as the
    reference refuses it (System::CheckInput)."""
    with pytest.raises(ValueError, match="lonely"):
        lat1(extra="lat : L : lowerbound : surface")


BOX27 = """
lat : box : gradients : 3
lat : box : n_layers_x : 8
lat : box : n_layers_y : 6
lat : box : n_layers_z : 10
lat : box : lattice_type : {lt}
lat : box : stencil_full : {full}
lat : box : lowerbound_x : periodic
lat : box : upperbound_x : periodic
lat : box : lowerbound_y : periodic
lat : box : upperbound_y : periodic
lat : box : lowerbound_z : periodic
lat : box : upperbound_z : periodic
mon : A : freedom : free
mon : W : freedom : free
mol : pol : composition : (A)5
mol : pol : freedom : free
mol : pol : phibulk : 0.1
mol : water : composition : (W)1
mol : water : freedom : solvent
"""


@pytest.mark.parametrize("lt,full", [("simple_cubic", "false"), ("hexagonal", "true"), ("hexagonal", "false")])
def test_3d_wide_stencil_symbols(lt, full):
    """The 26-neighbour 3-D stencils on a periodic box: a product of cosines
    is an eigenvector with eigenvalue (w0 + 2 wf sum c + 4 we sum cc + 8 wc ccc)
    -- simple cubic without stencil_full (0, 16, 4, 1)/152, hexagonal
    (8, 4, 2, 1)/40 (summing to 64/40, as in the reference), hexagonal
    without stencil_full (0, 4, 2, 1)/56 -- for the propagator and the site
    average alike."""
    o = Oracle(BOX27.format(lt=lt, full=full))
    MX, MY, MZ = 8, 6, 10
    w0, wf, we, wc, nrm = {("simple_cubic", "false"): (0, 16, 4, 1, 152), ("hexagonal", "true"): (8, 4, 2, 1, 40),
                           ("hexagonal", "false"): (0, 4, 2, 1, 56)}[(lt, full)]
    JX, JY = (MY + 2) * (MZ + 2), MZ + 2
    x, y, z = np.meshgrid(np.arange(1, MX + 1), np.arange(1, MY + 1), np.arange(1, MZ + 1), indexing="ij")
    idx = (x * JX + y * JY + z).ravel()
    for jx, jy, jz in ((1, 0, 0), (1, 2, 3), (4, 3, 5), (2, 1, 0)):
        kx, ky, kz = 2 * np.pi * jx / MX, 2 * np.pi * jy / MY, 2 * np.pi * jz / MZ
        v = np.zeros(o.M)
        v[idx] = (np.cos(kx * x + 0.2) * np.cos(ky * y + 0.5) * np.cos(kz * z - 0.3)).ravel()
        cx, cy, cz = np.cos(kx), np.cos(ky), np.cos(kz)
        ev = (w0 + 2 * wf * (cx + cy + cz) + 4 * we * (cx * cy + cx * cz + cy * cz) + 8 * wc * cx * cy * cz) / nrm
        for side in (False, True):
            np.testing.assert_allclose(o.apply(v, side=side)[idx], ev * v[idx], atol=1e-13)


def test_refined_2d_3d_block_sums():
    """Refined 2- and 3-gradient stencils (FJC 5, hexagonal) on mirror
    lattices: the 3-D block sum (weights 8/4/2/1 over 512) and the 2-D
    cylindrical one (2 LAMBDA / 2(fjc+2) with |dy| < fjc, LAMBDA / 2(fjc+2) at
    |dy| = fjc) keep a constant at every interior site; the 2-D planar block
    sum is normalised by 16 fjc^2 - 3 instead of 16 fjc^2 in the reference,
    so a constant grows by 64/61 there (fjc = 2)."""
    base = """
mon : A : freedom : free
mon : W : freedom : free
mol : pol : composition : (A)5
mol : pol : freedom : free
mol : pol : phibulk : 0.1
mol : water : composition : (W)1
mol : water : freedom : solvent
"""
    o3 = Oracle("lat : b : gradients : 3\nlat : b : n_layers_x : 3\nlat : b : n_layers_y : 4\n"
                "lat : b : n_layers_z : 3\nlat : b : lattice_type : hexagonal\nlat : b : FJC_choices : 5\n" + base)
    w = o3.interior(o3.apply(np.ones(o3.M)))
    np.testing.assert_allclose(w, 1.0, atol=1e-14)
    oc = Oracle("lat : c : gradients : 2\nlat : c : geometry : cylindrical\nlat : c : n_layers_x : 5\n"
                "lat : c : n_layers_y : 4\nlat : c : lattice_type : hexagonal\nlat : c : FJC_choices : 5\n" + base)
    np.testing.assert_allclose(oc.interior(oc.apply(np.ones(oc.M))), 1.0, atol=1e-14)
    op = Oracle("lat : p : gradients : 2\nlat : p : n_layers_x : 5\nlat : p : n_layers_y : 4\n"
                "lat : p : lattice_type : hexagonal\nlat : p : FJC_choices : 5\n" + base)
    np.testing.assert_allclose(op.interior(op.apply(np.ones(op.M))), 64.0 / 61.0, atol=1e-14)
