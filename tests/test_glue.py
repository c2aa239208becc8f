"""The reference-side binding (cucodeml_amd/glue/codeml_bsa_glue.c, linked into
the reference's codeml host program as oracle/_ref/codeml_bsa) refuses
control-file settings the engine does not compute, naming the key, before
any device call: another genetic code (icode, codeml.c:1498; tables at
tools.c:22), fixed branch lengths (fix_blength = 2, codeml.c:617-627) and a
clock.  Runs on the CPU: with icode = 0 the same run reaches the engine's
create call, which here fails for want of a GPU, so the refusal is what
stops the other runs."""
import os
import subprocess
import tempfile

import pytest

from . import fixtures

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "oracle", "_ref", "codeml_bsa")

pytestmark = pytest.mark.skipif(not os.path.exists(BIN), reason="oracle/_ref/codeml_bsa not "
                                "built (needs the reference sources at build time)")


def mito_safe(phylip):
    """cfg0's alignment with AGA / AGG (stops in the vertebrate mitochondrial
    code) replaced by CGA, so that codeml reads it under icode = 1 too."""
    head, *rows = phylip.strip().splitlines()
    out = [head]
    for r in rows:
        name, seq = r.split()
        cod = [seq[i:i + 3] for i in range(0, len(seq), 3)]
        out.append(f"{name}  " + "".join("CGA" if c in ("AGA", "AGG") else c for c in cod))
    return "\n".join(out) + "\n"


def run(settings, extra_start=()):
    fx, tree, _ = fixtures.load("cfg0")
    ctl = fixtures.CTL.format(fix_omega=0, omega=1.5)
    for k, v in settings.items():
        ctl = "\n".join(f"{k} = {v}" if line.split("=")[0].strip() == k else line
                        for line in ctl.splitlines()) + "\n"
    start = ([] if settings.get("fix_blength") == 2 else list(fx["ml"]["H1"]["start"][:-5])) \
        + list(fx["ml"]["H1"]["start"][-5:])
    with tempfile.TemporaryDirectory() as d:
        open(os.path.join(d, "seq.phy"), "w").write(mito_safe(fx["phylip"]))
        open(os.path.join(d, "tree.nwk"), "w").write(f"{fx['ns']} 1\n{fx['newick']}\n")
        open(os.path.join(d, "in.codeml"), "w").write(" ".join(f"{v:.12f}" for v in start) + "\n")
        open(os.path.join(d, "codeml.ctl"), "w").write(ctl)
        env = dict(os.environ, HIP_VISIBLE_DEVICES=os.environ.get("HIP_VISIBLE_DEVICES", ""))
        return subprocess.run([BIN, "codeml.ctl"], cwd=d, capture_output=True, text=True,
                              input="\n", timeout=120, env=env)


@pytest.mark.parametrize("key,value", [("icode", 1), ("fix_blength", 2), ("clock", 1)])
def test_refuses_settings_the_engine_does_not_compute(key, value):
    r = run({key: value})
    text = r.stdout + r.stderr
    assert r.returncode != 0
    if key == "clock":
        # codeml itself refuses a clock on an unrooted tree before the engine
        # is reached; either way nothing is computed
        assert "bsa_create" not in text
        return
    assert f"bsa: {key} = {value} is not offloaded" in text, text[-1500:]
    assert "bsa_create" not in text


def test_universal_code_reaches_the_engine():
    r = run({"icode": 0})
    text = r.stdout + r.stderr
    assert "is not offloaded" not in text
    if r.returncode != 0:
        # no GPU here: the engine's create call is what fails
        assert "bsa: bsa_create" in text, text[-1500:]
