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


def test_struct_layouts_match_header(pm):
    import ctypes as C
    assert C.sizeof(pm.BuildParams) == 11 * 4
    assert C.sizeof(pm.TraceParams) == 16
    # pm_tree: int + 9 ptrs + int + ptr + int + ptr + 3 ptrs + 4 ints (with padding)
    assert C.sizeof(pm.Tree) == 8 + 9 * 8 + 8 + 8 + 8 + 8 + 3 * 8 + 16
