"""Generate the output-format fixtures from the reference Namics program.

Each input below is a case of tests/golden/cases.py (or charged_cases.py)
with `kal` / `pro` / `vtk` / `vec` requests, DIIS to 1e-9 and, for some, a second start block and a final guess
file.  The reference (oracle/_ref/namics, built from its own sources) runs it
in a scratch directory; every file it writes to output/ is recorded as text.
Run here (where the reference binary is built); writes
tests/golden/output_files.json.

Usage: python tests/golden/make_output_fixtures.py
"""
import json
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

from tests.golden.cases import CASES  # noqa: E402
from tests.golden.charged_cases import CHARGED  # noqa: E402
from tests.golden.make_fixtures import REF  # noqa: E402

DIIS = "newton : isaac : method : DIIS\nnewton : isaac : tolerance : 1e-9\nnewton : isaac : iterationlimit : 3000\n"

KAL_1D = """output : kal : append : false
kal : sys : NN : free_energy
kal : sys : NN : grand_potential
kal : sys : NN : start
kal : newton : isaac : iterations
kal : newton : isaac : residual
kal : newton : isaac : method
kal : newton : isaac : m
kal : newton : isaac : delta_max
kal : mol : pol : GN
kal : mol : pol : n
kal : mol : pol : theta
kal : mol : pol : theta_exc
kal : mol : pol : phibulk
kal : mol : pol : Mu
kal : mol : pol : chainlength
kal : mol : pol : composition
kal : mol : pol : freedom
kal : mol : pol : phi1
kal : mol : pol : nonsense
kal : mol : water : Mu
kal : mon : A : theta
kal : mon : A : phibulk
kal : mon : A : freedom
kal : lat : flat : gradients
kal : lat : flat : volume
kal : lat : flat : lattice_type
kal : lat : flat : n_layers
"""

SEARCH_KAL = """output : kal : append : false
kal : sys : NN : grand_potential
kal : sys : NN : free_energy
kal : mol : pol : phibulk
kal : mol : pol : theta
kal : mol : pol : Mu
"""
DIIS10 = "newton : isaac : method : DIIS\nnewton : isaac : tolerance : 1e-10\nnewton : isaac : iterationlimit : 3000\n"

INPUTS = {
    # two start blocks: kal rows appended, one pro file per start, the
    # second start from the first one's fields, a guess file per start
    "scan1d": CASES["cfg1_planar_adsorption"]["text"] + KAL_1D + """output : pro : append : false
pro : mon : A : phi
pro : mon : W : phi
pro : mol : pol : phi
pro : mol : water : phi
sys : NN : final_guess : file
""" + DIIS + "start\nmon : A : chi_W : 0.6\nstart\n",
    "sph1d": CASES["sph1d_particle"]["text"] + """output : kal : append : false
kal : sys : NN : free_energy
kal : mol : * : theta
kal : mon : S : theta
kal : lat : L : volume
kal : lat : L : offset_first_layer
output : pro : append : false
pro : mon : * : phi
pro : mol : pol : phi
""" + DIIS + "start\n",
    "fjc5": CASES["fjc5_planar_adsorption"]["text"] + """output : kal : append : false
output : kal : header_separator : classic
kal : sys : NN : free_energy
kal : mol : pol : theta
kal : lat : flat : FJC_choices
output : pro : append : false
output : pro : write_bounds : true
pro : mon : A : phi
pro : mol : pol : phi
""" + DIIS + "start\n",
    "brush2d": CASES["brush2d_cyl"]["text"] + """output : kal : append : false
kal : sys : NN : free_energy
kal : sys : NN : grand_potential
kal : mol : brush : GN
kal : mol : brush : theta
kal : mon : X : theta
kal : lat : cyl : n_layers_x
kal : lat : cyl : n_layers_y
output : pro : append : false
pro : mon : A : phi
pro : mol : brush : phi
output : vtk : write_bounds : true
vtk : mon : A : phi
""" + DIIS + "start\n",
    # var scans: phibulk exponentially (2 points per decade), chi linearly
    # ((0.8 - 0.5) / 0.1 truncates to 2 in floating point: 3 points)
    "scan_phibulk": CASES["cfg1_planar_adsorption"]["text"] + """output : kal : append : false
kal : sys : NN : free_energy
kal : mol : pol : phibulk
kal : mol : pol : theta_exc
kal : newton : isaac : iterations
output : pro : append : false
pro : mol : pol : phi
var : mol-pol : scan : phibulk
var : mol-pol : scale : exponential
var : mol-pol : steps : 2
var : mol-pol : end_value : 0.1
""" + DIIS + "start\n",
    "scan_chi": CASES["sph1d_particle"]["text"] + """output : kal : append : false
kal : sys : NN : free_energy
kal : sys : NN : grand_potential
kal : mol : pol : theta
var : mon-A : scan : chi_W
var : mon-A : step : 0.1
var : mon-A : end_value : 0.7
sys : NN : final_guess : file
""" + DIIS + "start\n",
    # the polyelectrolyte's valence scanned linearly (a charged system;
    # the potential profile and its drop written too)
    "scan_valence": CHARGED["el1d_poly"]["text"] + """output : kal : append : false
kal : sys : NN : free_energy
kal : sys : NN : grand_potential
kal : sys : NN : Dpsi
kal : mol : pol : theta
kal : mol : Na : phibulk
kal : mon : A : valence
output : pro : append : false
pro : sys : NN : psi
pro : sys : NN : q
var : mon-A : scan : valence
var : mon-A : step : -0.2
var : mon-A : end_value : -0.6
""" + DIIS + "start\n",
    # vec files: every requested profile on one line per solve, appended
    # (two start blocks), a frozen segment's mask without its bounds set;
    # pos files: no object offers an integer array, so the file stays empty
    "vec1d": CASES["cfg1_planar_adsorption"]["text"] + """output : vec : append : false
vec : mon : S : phi
vec : mon : A : phi
vec : mol : * : phi
vec : mon : A : nonsense
output : pos : append : true
pos : mon : S : phi
""" + DIIS + "start\nmon : A : chi_W : 0.6\nstart\n",
    "vec_charged": CHARGED["el1d_poly"]["text"] + """output : vec : append : true
vec : sys : NN : psi
vec : sys : NN : q
vec : mon : A : phi
""" + DIIS + "start\n",
    "box3d": CASES["diblock3d_periodic"]["text"] + """output : kal : append : false
kal : sys : NN : free_energy
kal : mol : surf : GN
kal : mol : surf : theta
kal : mon : B : theta
output : pro : append : false
pro : mon : B : phi
pro : mol : water : phi
output : vtk : append : false
vtk : mol : surf : phi
""" + DIIS + "start\n",
    # var searches: the bulk volume fraction of the adsorbing polymer varied
    # until the grand potential reaches a set value, and (at another chi)
    # until the solvent's amount theta does (a search and a target on one
    # molecule give NaN results in the reference); each from the first
    # start block's solution
    "search_omega": CASES["cfg1_planar_adsorption"]["text"] + SEARCH_KAL + DIIS10 + """start
var : mol-pol : search : phibulk
var : sys-NN : grand_potential : -0.0025
start
""",
    "search_theta": CASES["cfg1_planar_adsorption"]["text"] + SEARCH_KAL + DIIS10 + """mon : A : chi_W : 0.6
start
var : mol-pol : search : phibulk
var : mol-water : theta : 98.4
start
""",
}


def main():
    if not os.path.exists(REF):
        sys.exit("build the reference first: make -C oracle ref")
    out = {}
    for name, text in INPUTS.items():
        with tempfile.TemporaryDirectory() as d:
            os.makedirs(os.path.join(d, "output"))
            with open(os.path.join(d, name + ".in"), "w") as f:
                f.write(text)
            r = subprocess.run([REF, name + ".in"], cwd=d, capture_output=True, text=True, timeout=600)
            if r.returncode != 0:
                raise RuntimeError(name + ": " + r.stdout[-2000:] + r.stderr[-2000:])
            files = {}
            for fn in sorted(os.listdir(os.path.join(d, "output"))):
                with open(os.path.join(d, "output", fn)) as f:
                    files[fn] = f.read()
        out[name] = {"input": text, "files": files}
        print(name, {k: len(v) for k, v in files.items()}, flush=True)
    with open(os.path.join(HERE, "output_files.json"), "w") as f:
        json.dump(out, f)


if __name__ == "__main__":
    main()
