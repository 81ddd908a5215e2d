"""Writes trans_args.json.gz: the ARGs of the reference's own transition
tests (test/data/test_trans/*.arg, 40 ARGs of 8 leaves, and
test/data/test_trans_switch/*.arg, 100 ARGs of 12 leaves; the files
test/test_hmm.py:117-190 reads), as data: per ARG its region and its nodes
(name, event, age, position, parents, children), in file order.

Run in a container that holds the reference checkout:
    python tests/golden/make_trans_fixtures.py [/root/reference]
"""
import gzip
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def read_arg_file(path):
    """The arglib text format: a 'start=.. end=..' line, a header line, then
    one node per line (name, event, age, pos, comma-separated parents and
    children)."""
    lines = open(path).read().splitlines()
    region = dict(kv.split("=") for kv in lines[0].split())
    nodes = []
    for line in lines[2:]:
        if not line.strip():
            continue
        f = line.split("\t")
        while len(f) < 6:
            f.append("")
        nodes.append(dict(name=f[0], event=f[1], age=float(f[2]), pos=float(f[3]),
                          parents=[x for x in f[4].split(",") if x],
                          children=[x for x in f[5].split(",") if x]))
    return dict(start=float(region["start"]), end=float(region["end"]), nodes=nodes)


def main(ref="/root/reference"):
    out = {}
    for name, n in (("test_trans", 40), ("test_trans_switch", 100)):
        d = os.path.join(ref, "test", "data", name)
        out[name] = [read_arg_file(os.path.join(d, "%d.arg" % i)) for i in range(n)]
    with gzip.open(os.path.join(HERE, "trans_args.json.gz"), "wt") as f:
        json.dump(out, f, separators=(",", ":"))
    print({k: len(v) for k, v in out.items()})


if __name__ == "__main__":
    main(*sys.argv[1:])
