"""Standard errors on the GPU (codeml.standard_errors: 2 np engine
evaluations of the per-site log f) against the reference's getSE = 1
printout for the same data and parameters, under H1 and H0, and through the
driver (control file with getSE = 1 -> "SEs for parameters" in out)."""
import gzip
import json
import os
import tempfile

import numpy as np
import pytest

from cucodeml_amd import codeml, engine
from . import fixtures

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_se(name):
    with gzip.open(os.path.join(GOLDEN, f"se_{name}.json.gz"), "rt") as f:
        return json.load(f)


@pytest.mark.parametrize("name", ["toy_amb", "cfg0"])
@pytest.mark.parametrize("hyp", ["H1", "H0"])
def test_standard_errors_match_reference(name, hyp):
    print(f"case: SE {name} {hyp}", flush=True)
    fx, tree, data = fixtures.load(name)
    g = load_se(name)[hyp]
    eng = engine.BranchSiteEngine(tree, data, fix_omega=hyp == "H0", omega_fix=1.0)
    try:
        se = codeml.standard_errors(eng, g["x"])
    finally:
        eng.close()
    ref = np.array(g["se"])
    print(f"  max rel {np.max(np.abs(se - ref) / ref):.2e}", flush=True)
    assert np.all(np.abs(se - ref) <= 2e-4 * ref + 1e-6)


def test_driver_prints_standard_errors():
    print("case: driver SE cfg0 H1 fixed parameters", flush=True)
    fx, _, _ = fixtures.load("cfg0")
    g = load_se("cfg0")["H1"]
    with tempfile.TemporaryDirectory() as d:
        open(os.path.join(d, "seq.phy"), "w").write(fx["phylip"])
        open(os.path.join(d, "tree.nwk"), "w").write(fx["newick"] + "\n")
        open(os.path.join(d, "in.codeml"), "w").write("-1 " + " ".join(f"{v:.12f}" for v in g["x"]))
        ctl = fixtures.CTL.format(fix_omega=0, omega=1.5).replace("getSE = 0", "getSE = 1")
        open(os.path.join(d, "codeml.ctl"), "w").write(ctl)
        codeml.run(os.path.join(d, "codeml.ctl"))
        out = open(os.path.join(d, "out")).read()
    line = out[out.index("SEs for parameters:"):].splitlines()[1]
    se = np.array([float(v) for v in line.split()])
    assert np.all(np.abs(se - np.array(g["se"])) <= 2e-4 * np.array(g["se"]) + 2e-6)
