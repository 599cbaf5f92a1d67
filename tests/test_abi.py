"""libpmkd.so loads and exports exactly what include/pmkd.h declares (no GPU
work: only the parameter-default functions, which touch no device, run)."""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "pmkd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_]+\s*\**\s*(pm_[a-z_0-9]+)\s*\(", text, re.M)))


def test_header_declares_package_exports(pm):
    assert declared_functions() == sorted(pm.EXPORTS)


def test_library_exports_every_declared_symbol(pm):
    lib = pm.load_library()
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_defaults_are_reference_constants(pm):
    p = pm.build_params("default")
    assert p.lower_algorithm == pm.LOWER_BITMAP          # TriangleMap.cu:46
    assert p.empty_space_splitting == 1                  # TriangleMapUpperCreator.cu:31
    assert abs(p.empty_space_threshold - 0.35) < 1e-7    # TriangleMapUpperCreator.cu:28
    assert p.split_method == pm.SPLIT_BOX                # TriangleMapUpperCreator.cu:32
    assert p.propagate_aabbs == 1                        # TriangleMap.cu:47
    assert p.traversal_cost == 24.0                      # TriangleMap.cu:48
    assert p.lower_size == 32                            # KDNode.h:21
    t = pm.trace_params()
    assert t.leaf_skipping == 1                          # IRayTracer.cu:33-34
    assert t.stack_size == 4                             # ShortStack.cu:25
    r = pm.build_params("reference")
    assert r.lower_algorithm == pm.LOWER_SAH


def test_struct_layouts_match_header(pm, tmp_path):
    """every ctypes struct field sits at the offset the C compiler gives it"""
    import ctypes as C
    import subprocess
    structs = {"pm_build_params": pm.BuildParams, "pm_trace_params": pm.TraceParams,
               "pm_tree": pm.Tree, "pm_stats": pm.Stats, "pm_tuning": pm.Tuning,
               "pm_graph_caps": pm.GraphCaps, "pm_light": pm.Light}
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "pmkd.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)])
    got = {}
    for line in subprocess.check_output([str(exe)], text=True).splitlines():
        cname, field, val = line.split()
        got[(cname, field)] = int(val)
    for cname, py in structs.items():
        assert got[(cname, "size")] == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)
