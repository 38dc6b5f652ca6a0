"""python -m vagen MODEL.va -o OUT.h --prefix NAME [-I DIR]... [-U MACRO]... [--inst P1,P2,...]"""
import argparse
import sys

from .emit import emit_cpp
from .parse import parse_module
from .pre import preprocess


def main(argv=None):
    ap = argparse.ArgumentParser(prog="vagen")
    ap.add_argument("model")
    ap.add_argument("-o", "--out", required=True)
    ap.add_argument("--prefix", required=True)
    ap.add_argument("-I", dest="incs", action="append", default=[])
    ap.add_argument("-U", dest="undef", action="append", default=[])
    ap.add_argument("--inst", default="", help="comma-separated instance parameters")
    a = ap.parse_args(argv)
    text = preprocess(a.model, include_dirs=a.incs, undefines=a.undef)
    mod = parse_module(text)
    inst = [x for x in a.inst.split(",") if x]
    src = emit_cpp(mod, inst, a.prefix)
    with open(a.out, "w") as f:
        f.write(src)
    print(f"vagen: {mod.name}: {len(mod.nodes)} nodes, {len(mod.params)} parameters, "
          f"{len(mod.vars)} variables -> {a.out}", file=sys.stderr)


if __name__ == "__main__":
    main()
