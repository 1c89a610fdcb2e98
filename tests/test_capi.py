"""The C-ABI library loads and exports every entry point include/*.h declares
(no compute call: this runs without a GPU)."""
import ctypes
import glob
import os
import re

from sparlab_amd import native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names.update(re.findall(r"\b(slx_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_every_declared_entry_point():
    assert os.path.exists(native.LIB_PATH), "build the library first (python sparlab_amd/build.py)"
    lib = ctypes.CDLL(native.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 30
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_every_declared_entry_point():
    assert declared_symbols() == set(native._SIGNATURES)


def test_binding_struct_layout_matches_header(tmp_path):
    # sizes and field offsets of the option / result structs as gcc lays them out
    import shutil
    import subprocess

    import pytest

    if not shutil.which("gcc"):
        pytest.skip("gcc not available")
    structs = {"slx_amg_options": native.AmgOptions, "slx_pcg_result": native.PcgResult,
               "slx_amg_stats": native.AmgStats, "slx_eval_options": native.EvalOptions,
               "slx_evaluation": native.Evaluation, "slx_oc_options": native.OcOptions,
               "slx_oc_step": native.OcStep, "slx_run_options": native.RunOptions,
               "slx_warmstart_options": native.WarmStartOptions, "slx_warmstart_stats": native.WarmStartStats,
               "slx_run_result": native.RunResult, "slx_projection_options": native.ProjectionOptions,
               "slx_modal_options": native.ModalOptions, "slx_modal_result": native.ModalResult}
    src = tmp_path / "layout.c"
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "sparlab_hip.h"', "int main(void) {"]
    for cname, cls in structs.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
    lines.append('printf("run.eval %zu\\n", offsetof(slx_run_options, eval));')
    for cname, cls in structs.items():  # every field's offset (ctypes names carry a trailing _ for C keywords)
        for fname, _ in cls._fields_:
            lines.append(f'printf("{cname}.{fname} %zu\\n", offsetof({cname}, {fname.rstrip("_")}));')
    lines.append("return 0; }")
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = dict(l.split() for l in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines())
    for cname, cls in structs.items():
        assert int(out[cname]) == ctypes.sizeof(cls), cname
    assert int(out["run.eval"]) == native.RunOptions.eval.offset
    for cname, cls in structs.items():
        for fname, _ in cls._fields_:
            assert int(out[f"{cname}.{fname}"]) == getattr(cls, fname).offset, f"{cname}.{fname}"
