"""The C-ABI library (libphx.so) loads without a GPU and exports every entry point declared in
include/phx.h; the ctypes structures match the header's layout. No compute calls here."""
import ctypes
import os
import re

import pytest

from phare_amd import phx

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "phx.h")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:phx_status|void\*|void|const char\*)\s+(phx_\w+)\s*\(",
                                 text, re.M)))


def test_header_declares_the_hot_path():
    names = declared()
    for n in ("phx_create", "phx_update_populations", "phx_faraday", "phx_ampere", "phx_ohm",
              "phx_advance", "phx_set_particles_aos", "phx_load_maxwellian"):
        assert n in names


@pytest.mark.skipif(not os.path.exists(phx.LIB_PATH), reason="libphx.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(phx.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(phx.EXPORTED) == declared()


def test_struct_layouts(tmp_path):
    """the ctypes mirrors have the size and field offsets the C compiler gives the header"""
    import shutil
    import subprocess

    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    src = tmp_path / "layout.c"
    src.write_text(
        '#include "phx.h"\n#include <stdio.h>\n#include <stddef.h>\n'
        'int main(void){printf("%zu %zu %zu %zu %zu\\n", sizeof(phx_config), sizeof(phx_patch),'
        ' offsetof(phx_config, Te), offsetof(phx_patch, level),'
        ' offsetof(phx_patch, nonlevel_ghost_upper));return 0;}\n')
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True,
                                          check=True).stdout.split()]
    assert got == [ctypes.sizeof(phx.Config), ctypes.sizeof(phx.Patch), phx.Config.Te.offset,
                   phx.Patch.level.offset, phx.Patch.nonlevel_ghost_upper.offset]


def test_particle_record_sizes():
    # Particle<dim> of the reference: 56 / 64 / 80 bytes
    assert [phx.particle_dtype(d).itemsize for d in (1, 2, 3)] == [56, 64, 80]
