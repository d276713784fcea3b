"""The reference-side binding (the ForceField subclass of the integration
notes at the repository root) compiles against the reference's own headers.

The block between the binding markers is written out as ff_dmd_gpu.h and
compiled with g++ -fsyntax-only next to src/ff.h, src/atoms.h, ... of the
reference (and the image's MPI headers, which they include):
  * once as MAPP::ff_style expands the style X-macro (src/MAPP.cpp:251-263),
    so the three style names resolve to constructible classes;
  * once with a function that creates each style through a ForceField
    pointer: that fails if a pure virtual of ForceField (energy_calc, calc_g,
    src/ff.h:21,28) is left unimplemented.
The reference's headers use strcmp / strcasecmp / memcpy / va_list without
including their headers, so <cstring>, <strings.h> and <cstdarg> are
force-included, as in the recipe that builds the reference under oracle/.
Skipped where the reference sources or the MPI headers are absent."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_SRC = "/root/reference/src"
MPI_INC = "/opt/conda/include"


def _binding():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    m = re.search(r"<!-- binding:begin -->\s*```cpp\n(.*?)```\s*<!-- binding:end -->", text, re.S)
    assert m, "no binding block between the binding markers"
    return m.group(1)


need_ref = pytest.mark.skipif(
    not (os.path.exists(os.path.join(REF_SRC, "ff.h")) and os.path.exists(os.path.join(MPI_INC, "mpi.h"))
         and shutil.which("g++")),
    reason="reference sources, MPI headers or g++ absent")

CHECK = r'''
#include <cstring>
#include "MAPP.h"
#include "ff_dmd_gpu.h"
using namespace MAPP_NS;
// the style table as MAPP::ff_style expands it
ForceField* pick(const char* name, MAPP* mapp) {
    ForceField* forcefield = NULL;
#define FF_Style
#define FFStyle(class_name,style_name) else if (strcmp(name, #style_name) == 0) forcefield = new class_name(mapp);
    if (0) {}
#include "ff_dmd_gpu.h"
#undef FF_Style
    return forcefield;
}
// every style is a complete ForceField: no pure virtual left
ForceField* make_all(MAPP* mapp) {
    ForceField* a = new ForceField_DMD_GPU(mapp);
    ForceField* b = new ForceField_DMD_I_GPU(mapp);
    ForceField* c = new ForceField_DMD_II_GPU(mapp);
    delete a;
    delete b;
    return c;
}
'''


@need_ref
def test_binding_compiles_against_reference_headers(tmp_path):
    (tmp_path / "ff_dmd_gpu.h").write_text(_binding())
    (tmp_path / "check.cpp").write_text(CHECK)
    cmd = ["g++", "-std=gnu++11", "-fsyntax-only", "-Wall", "-Wno-unused-variable",
           "-include", "cstring", "-include", "strings.h", "-include", "cstdarg",
           "-I", str(tmp_path), "-I", REF_SRC, "-I", os.path.join(ROOT, "include"), "-I", MPI_INC,
           str(tmp_path / "check.cpp")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    errors = [l for l in r.stderr.splitlines() if "ff_dmd_gpu.h" in l or "check.cpp" in l]
    assert r.returncode == 0, r.stderr[-4000:]
    assert not [l for l in errors if "warning" in l], "\n".join(errors)


def test_binding_uses_only_declared_entry_points():
    """every mapp_dmd_* the binding calls is declared in include/mapp_dmd.h"""
    header = open(os.path.join(ROOT, "include", "mapp_dmd.h")).read()
    declared = set(re.findall(r"\b(mapp_dmd_\w+)\s*\(", header))
    used = set(re.findall(r"\b(mapp_dmd_\w+)\s*\(", _binding()))
    assert used and used <= declared, used - declared
