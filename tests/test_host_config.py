"""The C host's starrynight.cfg reader (starrynight_amd/csrc/host/sn_config.c):
libconfig grammar subset and libconfig's no-auto-conversion typing, which the
reference's load_config relies on (src/starrynight-config.c:98-182).

A small C harness is compiled with gcc against the reader and asked for
typed lookups; no GPU and no HIP are involved."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "starrynight_amd", "csrc", "host")

HARNESS = r"""
#include <stdio.h>
#include <string.h>
#include "sn_config.h"
/* usage: harness file (type:path)...   prints one line per lookup */
int main(int argc, char **argv) {
    snc_config cf;
    if (snc_read_file(&cf, argv[1])) { printf("ERROR %d %s\n", cf.error_line, cf.error); return 0; }
    for (int i = 2; i < argc; i++) {
        char t = argv[i][0]; const char *p = argv[i] + 2;
        if (t == 'i') { int v; if (snc_lookup_int(&cf, p, &v)) printf("%s=%d\n", p, v); else printf("%s=-\n", p); }
        if (t == 'f') { double v; if (snc_lookup_float(&cf, p, &v)) printf("%s=%.17g\n", p, v); else printf("%s=-\n", p); }
        if (t == 'b') { int v; if (snc_lookup_bool(&cf, p, &v)) printf("%s=%d\n", p, v); else printf("%s=-\n", p); }
        if (t == 's') { const char *v; if (snc_lookup_string(&cf, p, &v)) printf("%s=%s\n", p, v); else printf("%s=-\n", p); }
        if (t == 'a') { const snc_node *n = snc_lookup(&cf, p); int L = snc_length(n); printf("%s=", p);
                        for (int k = 0; k < L; k++) printf("%s%.17g", k ? "," : "", snc_float_elem(n, k)); printf("\n"); }
    }
    snc_destroy(&cf);
    return 0;
}
"""


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    d = tmp_path_factory.mktemp("cfg")
    src = d / "h.c"
    src.write_text(HARNESS)
    exe = d / "h"
    subprocess.check_call(["gcc", "-O1", "-std=gnu11", "-Wall", "-I", HOST, "-o", str(exe), str(src),
                           os.path.join(HOST, "sn_config.c")])

    def run(text, *queries):
        cfg = d / "t.cfg"
        cfg.write_text(text)
        out = subprocess.check_output([str(exe), str(cfg), *queries], text=True)
        if out.startswith("ERROR"):
            return {"error": out}
        return dict(l.split("=", 1) for l in out.strip().splitlines())
    return run


CFG = """
# a run file in the reference's style
LOGFILE="run.log"
T: 250 # kelvin
X=16
Y=12 ; Z = 1
InitialLattice = "ferro" "electric"
Efield:
{
 x=0.0   # comment
 y=-0.05
 z=1e-2
}
DipoleCutOff: 4
ConstrainToX: TRUE
K: 0.5
Dipoles    = [ 1.0, 0.5, 0.0];
Prevalence = [ 0.25, 0.75, 0.0];
CageStrain = 1.5
MCMegaSteps: 0x14
MCEqmSteps: 5L
MCMoves: 200.0
// C++ comment
/* block
   comment */
Tile = [8, 8, 6];
"""


def test_reference_style_keys(harness):
    v = harness(CFG, "s:LOGFILE", "i:T", "i:X", "i:Y", "i:Z", "s:InitialLattice", "f:Efield.x", "f:Efield.y",
                "f:Efield.z", "i:DipoleCutOff", "b:ConstrainToX", "f:K", "a:Dipoles", "a:Prevalence",
                "f:CageStrain", "i:MCMegaSteps", "i:MCEqmSteps", "f:MCMoves", "a:Tile")
    assert v["LOGFILE"] == "run.log" and v["T"] == "250"
    assert (v["X"], v["Y"], v["Z"]) == ("16", "12", "1")
    assert v["InitialLattice"] == "ferroelectric"          # adjacent literals concatenate
    assert float(v["Efield.y"]) == -0.05 and float(v["Efield.z"]) == 0.01 and float(v["Efield.x"]) == 0.0
    assert v["DipoleCutOff"] == "4" and v["ConstrainToX"] == "1" and float(v["K"]) == 0.5
    assert v["Dipoles"] == "1,0.5,0" and v["Prevalence"] == "0.25,0.75,0"
    assert float(v["CageStrain"]) == 1.5
    assert v["MCMegaSteps"] == "20" and v["MCEqmSteps"] == "5" and float(v["MCMoves"]) == 200.0
    assert v["Tile"] == "0,0,0"   # integer elements read as float give 0, as libconfig's get_float_elem


def test_no_auto_conversion(harness):
    """libconfig without CONFIG_OPTION_AUTOCONVERT: an int lookup of a float
    setting (and the reverse) fails and the caller keeps its default."""
    v = harness("A = 3\nB = 3.0\nC = true\n", "f:A", "i:B", "i:C", "b:A", "i:A", "f:B", "i:missing")
    assert v["A"] == "3" and v["B"] == "3" and v["C"] == "-" and v["missing"] == "-"


def test_syntax_error_reports_line(harness):
    v = harness("X = 4\nY = [1, 2\nZ = 3\n", "i:X")
    assert "error" in v
    assert any(f"ERROR {n} " in v["error"] for n in (2, 3))


def test_reads_the_reference_cfg_file(harness):
    ref = "/root/reference/starrynight.cfg"
    if not os.path.exists(ref):
        pytest.skip("reference tree not mounted")
    v = harness(open(ref).read(), "i:T", "i:X", "i:Z", "s:InitialLattice", "i:DipoleCutOff", "b:ConstrainToX",
                "a:Dipoles", "f:CageStrain", "i:MCMegaSteps", "f:MCMoves", "f:Efield.x")
    assert "error" not in v
    assert v["T"] == "300" and v["X"] == "20" and v["Z"] == "28" and v["InitialLattice"] == "antiferro_wall"
    assert v["DipoleCutOff"] == "3" and v["ConstrainToX"] == "0" and v["Dipoles"] == "1,0,0"
    assert float(v["CageStrain"]) == 1.0 and v["MCMegaSteps"] == "20" and float(v["MCMoves"]) == 200.0
