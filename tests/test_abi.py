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
