"""Writes the tree/chain network fixtures used as inputs by the tests and the
benchmark, by running the reference's own generator (scripts/makenetwork.py)
in the build container.  The outputs are plain .netw text files; the generator
itself is not copied.  Run from the repository root:

    python tests/golden/make_networks.py /root/reference
"""
import contextlib
import importlib.util
import io
import os
import sys


def main(refroot):
    spec = importlib.util.spec_from_file_location(
        "makenetwork", os.path.join(refroot, "scripts", "makenetwork.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    outdir = os.path.join(os.path.dirname(__file__), "networks")
    os.makedirs(outdir, exist_ok=True)
    jobs = [(n, False) for n in list(range(6, 33)) + [36, 50]]
    jobs += [(n, True) for n in list(range(4, 33)) + [36, 50]]
    for n, dmrg in jobs:
        net = mod.network(sites=n, isDMRG=dmrg)
        buf = io.StringIO()
        with contextlib.redirect_stdout(buf):
            net.printnetworkfile()
        name = "%d_%s.netw" % (n, "DMRG" if dmrg else "T3NS")
        with open(os.path.join(outdir, name), "w") as f:
            f.write(buf.getvalue())


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
