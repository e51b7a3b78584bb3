"""The Python binding's ctypes structures have the C layout of include/gsj.h: the same size and the same offset for
every field (compiled with gcc from the header, no GPU needed). A trailing field added to a C structure and missed in
the binding would make the library read past the binding's structure."""
import os
import shutil
import subprocess

import pytest

import gsj

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
STRUCTS = [(gsj.JoinOpts, "gsj_join_opts"), (gsj.JoinStats, "gsj_join_stats"), (gsj.LayerInfo, "gsj_layer_info"),
           (gsj.Request, "gsj_request"), (gsj.RequestTimes, "gsj_request_times"), (gsj.CopyStats, "gsj_copy_stats")]


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_ctypes_structures_match_the_header(tmp_path):
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{os.path.join(ROOT, "include", "gsj.h")}"', "int main(void) {"]
    for cls, name in STRUCTS:
        lines.append(f'  printf("{name} size %zu\\n", sizeof({name}));')
        for f in cls._fields_:
            lines.append(f'  printf("{name} {f[0]} %zu\\n", offsetof({name}, {f[0]}));')
    lines.append("  return 0;\n}\n")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-o", str(exe), str(src)], check=True, capture_output=True)
    got = dict(l.rsplit(" ", 1) for l in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines())
    for cls, name in STRUCTS:
        assert int(got[f"{name} size"]) == gsj.C.sizeof(cls), name
        for f in cls._fields_:
            assert int(got[f"{name} {f[0]}"]) == getattr(cls, f[0]).offset, (name, f[0])
