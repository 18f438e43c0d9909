"""CPU checks of the C-ABI library: it loads and exports every entry point the header
declares (no GPU call is made)."""
import ctypes
import os
import re

from openfpm_data_amd import _capi as C

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "sparsegrid_gpu.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(sg_[a-z_0-9]+)\s*\(", txt)))


def test_header_matches_binding():
    assert _declared() == C.exported_symbols()


def test_library_exports_every_symbol():
    lib = ctypes.CDLL(C.LIB_PATH)
    for name in _declared():
        assert hasattr(lib, name), name


def test_stencil_desc_layout(tmp_path):
    """The ctypes mirror of sg_stencil_desc has the C compiler's size and offsets."""
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "sparsegrid_gpu.h"\n'
        'int main(void){printf("%zu %zu %zu %zu\\n", sizeof(sg_stencil_desc),'
        ' offsetof(sg_stencil_desc, use_box), offsetof(sg_stencil_desc, coef),'
        ' offsetof(sg_stencil_desc, icoef)); return 0;}\n')
    exe = tmp_path / "layout"
    import subprocess
    subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)])
    size, ub, co, ic = map(int, subprocess.check_output([str(exe)]).split())
    assert ctypes.sizeof(C.StencilDesc) == size
    assert C.StencilDesc.use_box.offset == ub
    assert C.StencilDesc.coef.offset == co
    assert C.StencilDesc.icoef.offset == ic


def test_kernel_view_layout(tmp_path):
    """The ctypes mirror of sg_kernel_view has the C compiler's size and offsets."""
    src = tmp_path / "kv.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "sparsegrid_gpu.h"\n'
        'int main(void){printf("%zu %zu %zu %zu %zu\\n", sizeof(sg_kernel_view),'
        ' offsetof(sg_kernel_view, nn), offsetof(sg_kernel_view, nchunks),'
        ' offsetof(sg_kernel_view, nnn), offsetof(sg_kernel_view, nprops)); return 0;}\n')
    exe = tmp_path / "kv"
    import subprocess
    subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)])
    size, nn, nch, nnn, npr = map(int, subprocess.check_output([str(exe)]).split())
    assert ctypes.sizeof(C.KernelView) == size
    assert C.KernelView.nn.offset == nn
    assert C.KernelView.nchunks.offset == nch
    assert C.KernelView.nnn.offset == nnn
    assert C.KernelView.nprops.offset == npr
