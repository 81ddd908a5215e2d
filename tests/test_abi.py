"""The C-ABI library loads and exports every function its header declares
(no compute call: this runs without a GPU)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "argweaver_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    names = set()
    for m in re.finditer(r"([A-Za-z_][A-Za-z0-9_]*)\s*\(", text):
        name = m.group(1)
        prefix = text[max(0, m.start() - 80):m.start()]
        if name in ("sizeof", "if", "for", "while", "return"):
            continue
        # a declaration: preceded by a type on the same statement
        if re.search(r"[A-Za-z_0-9\*]\s*\**\s*$", prefix.split(";")[-1].split("}")[-1]):
            names.add(name)
    return sorted(names)


def test_header_symbols_exported():
    from argweaver_amd import load_library
    lib = load_library()
    names = declared_functions()
    assert len(names) >= 30, names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_tree_arrays_round_trip_and_block_extent():
    """Host-side ARG handles (no GPU): the int** row tables the binding
    builds carry every row, an ARG rebuilt from its arrays exports the same
    arrays, and the columns covered by the blocks equal the summed block
    lengths."""
    import numpy as np
    import argweaver_amd.argweaverc as A
    times = np.array(A.get_time_points(20, 200e3))
    pops = np.full(20, 1e4)
    tr, _ = A.simulate(12, 30000, times, pops, 1.6e-8, 1.8e-8, 5)
    P, Ag, S, bl = A.trees_arrays(tr)
    assert P.shape[0] == len(bl) > 1
    assert A.seqlen_of(tr) == int(bl.sum()) == 30000
    tr2 = A.new_trees(P, Ag, S, bl)
    for x, y in zip((P, Ag, S, bl), A.trees_arrays(tr2)):
        assert np.array_equal(x, y)
    A.delete_local_trees(tr2)
    A.delete_local_trees(tr)


# internal-branch resampling entry points of the reference's binding that
# this library does not export yet
NOT_YET_EXPORTED = set()


def test_reference_binding_names_exported():
    """Every C entry point the reference's ctypes binding binds by name
    (argweaver/argweaverc.py:29-343, listed in tests/golden/
    argweaverc_exports.txt) is exported, save the listed internal-branch ones."""
    from argweaver_amd import load_library
    lib = load_library()
    names = open(os.path.join(ROOT, "tests", "golden", "argweaverc_exports.txt")).read().split()
    assert len(names) == 57
    missing = {n for n in names if not hasattr(lib, n)}
    assert missing == NOT_YET_EXPORTED
