"""C ABI: the library loads without a GPU and exports every function that
include/*.h declares; the ctypes structs match the header's layout."""
import ctypes
import glob
import os
import re

import pytest

from monte_carlompi_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, 'include', '*.h')):
        text = re.sub(r'/\*.*?\*/', '', open(h).read(), flags=re.S)
        names |= set(re.findall(r'^\s*(?:const\s+)?\w+\s*\*?\s*(mcs_\w+)\s*\(', text, flags=re.M))
    return names


def test_header_declares_the_boundary():
    names = declared_functions()
    for n in ('mcs_create', 'mcs_set_cases', 'mcs_walk', 'mcs_read_records', 'mcs_read_tallies'):
        assert n in names


def test_library_exports_every_declared_function():
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in sorted(declared_functions()) if not hasattr(L, n)]
    assert not missing, missing
    assert set(declared_functions()) <= set(_lib.EXPORTS) | set()


def test_struct_sizes_match_header():
    assert ctypes.sizeof(_lib.McsCase) == 56
    assert ctypes.sizeof(_lib.McsBand) == 64
    assert ctypes.sizeof(_lib.McsWalkArgs) == 56
    assert ctypes.sizeof(_lib.McsLaunch) == 20
    assert ctypes.sizeof(_lib.McsPolarCase) == 40


def test_struct_layouts_match_the_c_compiler(tmp_path):
    """sizeof / offsetof of every ABI struct from the C compiler on
    include/mcsnow.h equal the ctypes mirrors."""
    import shutil
    import subprocess
    if shutil.which('gcc') is None:
        pytest.skip('no C compiler')
    structs = {'mcs_case': _lib.McsCase, 'mcs_band': _lib.McsBand, 'mcs_walk_args': _lib.McsWalkArgs,
               'mcs_launch': _lib.McsLaunch, 'mcs_records': _lib.McsRecords, 'mcs_stats': _lib.McsStats,
               'mcs_polar_case': _lib.McsPolarCase, 'mcs_dev_out': _lib.McsDevOut}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "mcsnow.h"', 'int main(void) {']
    for cname, cls in structs.items():
        lines.append('printf("%s %%zu\\n", sizeof(%s));' % (cname, cname))
        for f, _ in cls._fields_:
            cf = 'async' if f == 'async_' else f
            lines.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (cname, f, cname, cf))
    lines.append('return 0; }')
    src = tmp_path / 'layout.c'
    src.write_text('\n'.join(lines))
    exe = tmp_path / 'layout'
    subprocess.check_call(['gcc', '-I', os.path.join(ROOT, 'include'), str(src), '-o', str(exe)])
    got = dict(l.split() for l in subprocess.check_output([str(exe)]).decode().splitlines())
    for cname, cls in structs.items():
        assert int(got[cname]) == ctypes.sizeof(cls), cname
        for f, _ in cls._fields_:
            assert int(got['%s.%s' % (cname, f)]) == getattr(cls, f).offset, (cname, f)
    assert ctypes.sizeof(_lib.McsRecords) == 64
    assert ctypes.sizeof(_lib.McsStats) == 40
    assert ctypes.sizeof(_lib.McsDevOut) == 64


def test_header_constants_match_the_bindings():
    text = open(os.path.join(ROOT, 'include', 'mcsnow.h')).read()
    assert int(re.search(r'#define MCS_NOALLOC_MAX_CASES (\d+)', text).group(1)) == _lib.NOALLOC_MAX_CASES
    assert int(re.search(r'#define MCS_TALLY_F64 (\d+)', text).group(1)) == _lib.TALLY_F64
    assert int(re.search(r'#define MCS_TALLY_U64 (\d+)', text).group(1)) == _lib.TALLY_U64


def test_create_without_gpu_fails_cleanly():
    import torch
    if torch.cuda.is_available():
        pytest.skip('a GPU is present')
    ctx = ctypes.c_void_p()
    rc = _lib.lib().mcs_create(0, ctypes.byref(ctx))
    assert rc != 0 and not ctx.value
    assert _lib.lib().mcs_destroy(None) == -1
