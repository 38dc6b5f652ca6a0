"""vagen — Verilog-A compact model to C++/HIP translator.

The reference compiles Verilog-A models through VerilogAParser.jl and
src/vasim.jl (make_spice_device, src/vasim.jl:663-875) into Julia functions
whose derivatives DAECompiler takes by automatic differentiation.  vagen does
the same job for the GPU path: it preprocesses and parses a model file and
emits one templated C++ function (host and device) that evaluates every
branch contribution, split into resistive currents and charges (the
arguments of ddt), with the scalar type as a template parameter so that a
dual-number type yields all partial derivatives with respect to the node
voltages in one pass.

    python -m vagen MODEL.va -o OUT.h [-I include_dir ...] [-U MACRO ...]
"""
from .pre import preprocess
from .parse import parse_module
from .emit import emit_cpp

__all__ = ["preprocess", "parse_module", "emit_cpp"]
