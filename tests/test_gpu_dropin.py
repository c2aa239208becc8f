"""Drop-in check: the reference's own codeml host program (codeml.c, tools.c
built with its CUDA hooks on) linked with the engine through the binding in
cucodeml_amd/glue/ in place of cuda-codeml.cu (oracle/_ref/codeml_bsa, built
by oracle/Makefile).  Its maximum-likelihood fits of H1 and H0 from the
seeded starts must match the reference CPU codeml's (golden fixture) to 1e-4
in lnL and every MLE, with ming2's finite-difference gradient (the
reference's CUDA configuration) and with the engine's adjoint gradient in
ming2's dfun slot."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from . import fixtures

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "oracle", "_ref", "codeml_bsa")


def run(fx, start, fix_omega, omega, env_extra):
    with tempfile.TemporaryDirectory() as d:
        open(os.path.join(d, "seq.phy"), "w").write(fx["phylip"])
        open(os.path.join(d, "tree.nwk"), "w").write(f"{fx['ns']} 1\n{fx['newick']}\n")
        open(os.path.join(d, "in.codeml"), "w").write(" ".join(f"{v:.12f}" for v in start) + "\n")
        open(os.path.join(d, "codeml.ctl"), "w").write(fixtures.CTL.format(fix_omega=fix_omega,
                                                                           omega=omega))
        env = dict(os.environ, **env_extra)
        r = subprocess.run([BIN, "codeml.ctl"], cwd=d, capture_output=True, text=True,
                           input="\n", timeout=240, env=env)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        return fixtures.parse_out(open(os.path.join(d, "out")).read())


@pytest.mark.skipif(not os.path.exists(BIN), reason="oracle/_ref/codeml_bsa not built "
                    "(needs the reference sources at build time)")
@pytest.mark.parametrize("dfun", ["fd", "adjoint"])
@pytest.mark.parametrize("hyp", ["H1", "H0"])
def test_reference_codeml_with_engine_matches_cpu_codeml(hyp, dfun):
    print(f"case: drop-in {hyp} {dfun}", flush=True)
    fx, _, _ = fixtures.load("cfg0")
    ml = fx["ml"][hyp]
    fo, om = (0, 1.5) if hyp == "H1" else (1, 1.0)
    lnl, npar, x = run(fx, ml["start"], fo, om, {"BSA_DFUN": "1" if dfun == "adjoint" else "0"})
    print(f"  lnL {lnl} vs cpu {ml['mle_lnL']}; max |dx| "
          f"{np.max(np.abs(np.array(x) - np.array(ml['mle_x']))):.2e}", flush=True)
    assert abs(lnl - ml["mle_lnL"]) <= 1e-4
    assert len(x) == len(ml["mle_x"])
    assert np.max(np.abs(np.array(x) - np.array(ml["mle_x"]))) <= 1e-4 + 1e-6
