"""Writes argweaverc_exports.txt: the C entry points the reference's ctypes
binding (argweaver/argweaverc.py:29-343) binds by name, one per line.

Run in a container that holds the reference checkout:
    python tests/golden/make_abi_names.py [/root/reference]
"""
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def main(ref="/root/reference"):
    text = open(os.path.join(ref, "argweaver", "argweaverc.py")).read()
    names = []
    for line in text.splitlines():
        if line.lstrip().startswith("#"):
            continue   # bindings commented out in the reference are not bound
        m = re.search(r'argweaverclib,\s*"(\w+)"', line)
        if m and m.group(1) not in names:
            names.append(m.group(1))
    with open(os.path.join(HERE, "argweaverc_exports.txt"), "w") as f:
        f.write("\n".join(names) + "\n")
    print(len(names), "names")


if __name__ == "__main__":
    main(*sys.argv[1:])
