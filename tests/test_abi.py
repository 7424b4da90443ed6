"""The C-ABI boundary -- CPU only (no compute calls).

* libaar_hip.so exports every function include/*.h declares;
* the POISSON descriptor the library is built with has the layout callers
  get from include/system.h with MPICH's <mpi.h>;
* the reference driver compiles against include/ unchanged (where the
  reference tree is present).
"""
import os
import re
import subprocess

import pytest

from conftest import REPO

INCLUDE = os.path.join(REPO, "include")
LIB = os.path.join(REPO, "linear-solver_amd", "libaar_hip.so")
MPI_PREFIX = "/opt/conda"


def declared_functions():
    names = set()
    decl = re.compile(r"^\s*(?:const\s+)?(?:void|int|double|char)\s*\*?\s*([A-Za-z_][A-Za-z0-9_]*)\s*\(",
                      re.M)
    for h in os.listdir(INCLUDE):
        if h.endswith(".h"):
            with open(os.path.join(INCLUDE, h)) as fp:
                txt = re.sub(r"/\*.*?\*/", "", fp.read(), flags=re.S)
            names.update(decl.findall(txt))
    return names


def exported_functions():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True,
                         check=True).stdout
    return {ln.split()[-1] for ln in out.splitlines() if " T " in ln}


@pytest.fixture(scope="module")
def built():
    if not os.path.exists(LIB):
        pytest.skip("libaar_hip.so not built")
    return LIB


def test_exports_every_declared_symbol(built):
    from aar_hip._lib import REFERENCE_SYMBOLS, EXTENSION_SYMBOLS
    decl = declared_functions()
    assert set(REFERENCE_SYMBOLS) | set(EXTENSION_SYMBOLS) == decl
    missing = decl - exported_functions()
    assert not missing, f"declared but not exported: {sorted(missing)}"


def test_library_loads_without_gpu(built):
    from aar_hip import _lib
    L = _lib.load_for_symbols()
    for name in _lib.REFERENCE_SYMBOLS + _lib.EXTENSION_SYMBOLS:
        assert hasattr(L, name)


LAYOUT_SRC = r"""
#include <stdio.h>
#include <stddef.h>
#include HEADER
#define F(f) printf(#f " %zu\n", offsetof(POISSON, f));
int main(void) {
  printf("sizeof %zu\n", sizeof(POISSON));
  F(FDn) F(ssize) F(psize) F(np) F(coords) F(rem) F(send_neighs) F(rec_neighs) F(send_counts)
  F(rec_counts) F(send_layers) F(rec_layers) F(scounts) F(sdispls) F(rcounts) F(rdispls) F(n_in)
  F(n_out) F(phi) F(rhs) F(Lap_phi) F(solver_tol) F(solver_maxiter) F(m) F(p) F(omega) F(beta)
  F(coeff_lap) F(comm_laplacian) F(cart)
  return 0;
}
"""


def _layout(tmp_path, tag, compiler, header, incs):
    src = tmp_path / f"lay_{tag}.c"
    exe = tmp_path / f"lay_{tag}"
    src.write_text(LAYOUT_SRC.replace("HEADER", f'"{header}"'))
    lang = ["-x", "c++"] if compiler == "g++" else []
    subprocess.run([compiler] + lang + [str(src), "-o", str(exe)] + [f"-I{i}" for i in incs],
                   check=True, capture_output=True)
    return subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout


def test_poisson_layout_matches_mpich(tmp_path):
    if not os.path.exists(os.path.join(MPI_PREFIX, "include", "mpi.h")):
        pytest.skip("MPICH headers not present")
    caller = _layout(tmp_path, "caller", "gcc", "system.h", [INCLUDE, MPI_PREFIX + "/include"])
    lib = _layout(tmp_path, "lib", "g++", "poisson_abi.h",
                  [os.path.join(REPO, "linear-solver_amd", "csrc")])
    assert caller == lib


def test_reference_driver_compiles_against_include(tmp_path):
    ref_main = "/root/reference/MPI_REAL/main.c"
    if not os.path.exists(ref_main) or not os.path.exists(os.path.join(MPI_PREFIX, "include", "mpi.h")):
        pytest.skip("reference tree or MPICH not present")
    subprocess.run(["gcc", "-c", ref_main, "-o", str(tmp_path / "main.o"), f"-I{INCLUDE}",
                    f"-I{MPI_PREFIX}/include"], check=True, capture_output=True)


def test_dropin_wrong_argc_matches_reference():
    """The reference driver (MPI_REAL/main.c) relinked with libaar_hip and
    given a wrong argument count: CheckInputs (system.c:37-39) prints the
    reference's banner and "Wrong inputs" and ends with its exit status, as
    the reference build does (tests/golden/golden_edge.json).  No GPU call
    happens before CheckInputs."""
    import json
    exe = os.path.join(REPO, "oracle", "_ref", "main_on_libaar")
    if not os.path.exists(exe):
        pytest.skip("drop-in driver not built (make -C oracle dropin)")
    with open(os.path.join(REPO, "tests", "golden", "golden_edge.json")) as fp:
        ref = json.load(fp)["argc"]
    for argc, r in ref.items():
        out = subprocess.run([exe] + r["args"], capture_output=True, text=True, timeout=120)
        got = re.sub(r"Starting time: .*", "Starting time: <masked>", out.stdout)
        assert out.returncode == r["rc"], argc
        assert got == r["stdout"], argc

