"""The C ABI's structs (include/dada2_amd.h) against their ctypes mirrors in
the package: size and every field's offset, from a probe compiled with gcc
against the header.  No GPU."""
import shutil
import subprocess

import pytest

PAIRS = [  # (C struct, ctypes class name)
    ("d2g_align_params", "AlignParams"),
    ("d2g_vec_scores", "VecScores"),
    ("d2g_stats", "Stats"),
    ("d2g_dada_params", "DadaParams"),
    ("d2g_cluster_info", "ClusterInfo"),
    ("d2g_dada_stats", "DadaStats"),
    ("d2g_bimera_params", "BimeraParams"),
    ("d2g_pair_stats", "PairStats"),
    ("d2g_pairs_dev", "PairsDev"),
    ("d2g_comparison", "Comparison"),
]


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_struct_layouts_match(pkg, tmp_path):
    from pathlib import Path
    inc = Path(__file__).resolve().parents[1] / "include"
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "dada2_amd.h"',
             "int main(void) {"]
    for cname, pyname in PAIRS:
        cls = getattr(pkg, pyname)
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for field, _ in cls._fields_:
            lines.append(f'  printf("{cname} {field} %zu\\n", offsetof({cname}, {field}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-std=c11", "-I", str(inc), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    got = {}
    for line in out.splitlines():
        cname, key, val = line.split()
        got[(cname, key)] = int(val)
    for cname, pyname in PAIRS:
        cls = getattr(pkg, pyname)
        import ctypes as C
        assert got[(cname, "size")] == C.sizeof(cls), cname
        for field, _ in cls._fields_:
            assert got[(cname, field)] == getattr(cls, field).offset, (cname, field)
