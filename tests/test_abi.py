"""The C ABI: the library loads and exports every entry point include/jellyfysh_amd.h
declares (no device calls here)."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "jellyfysh_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ecmc_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    from jellyfysh_amd import _lib
    L = _lib.lib()
    names = declared_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) == set(_lib.exported_symbols())


def test_structs_match_header_sizes(tmp_path):
    import ctypes
    import shutil
    import subprocess
    from jellyfysh_amd import _lib
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler to read the header's struct sizes")
    src = tmp_path / "sizes.c"
    src.write_text('#include "jellyfysh_amd.h"\n#include <stdio.h>\n#include <stddef.h>\n'
                   'int main(void){printf("%zu %zu %zu %zu %zu %zu %zu\\n", sizeof(ecmc_tagger_desc),'
                   ' sizeof(ecmc_system_desc), sizeof(ecmc_event), offsetof(ecmc_system_desc, taggers),'
                   ' sizeof(ecmc_energy_term), sizeof(ecmc_batch_desc), sizeof(ecmc_batch_event));return 0;}\n')
    exe = tmp_path / "sizes"
    subprocess.run([cc, "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    sizes = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    from jellyfysh_amd.batch import BATCH_EVENT_DTYPE, BatchDesc
    assert sizes == [ctypes.sizeof(_lib.TaggerDesc), ctypes.sizeof(_lib.SystemDesc), _lib.EVENT_DTYPE.itemsize,
                     _lib.SystemDesc.taggers.offset, ctypes.sizeof(_lib.EnergyTerm), ctypes.sizeof(BatchDesc),
                     BATCH_EVENT_DTYPE.itemsize]


def test_device_count_without_gpu_is_not_an_error():
    from jellyfysh_amd import _lib
    n = _lib.lib().ecmc_device_count()
    assert n >= 0
