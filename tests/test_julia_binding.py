"""Static checks of the Julia binding (maxwellfdm.jl_amd/julia/MaxwellFDFDHIP)
against the C ABI, since Julia itself is not available to run it: every
symbol it ccalls is declared in include/mfdfd.h and exported by libmfdfd,
its structs list the C structs' fields in the same order with the same
widths, and every function the reference exports (model.jl:23-28) is
exported by the binding."""
import ctypes as C
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
JL = os.path.join(ROOT, "maxwellfdm.jl_amd", "julia", "MaxwellFDFDHIP", "src", "MaxwellFDFDHIP.jl")
HEADER = os.path.join(ROOT, "include", "mfdfd.h")

JULIA_WIDTH = {"Int64": 8, "Int32": 4, "Float64": 8, "ComplexF64": 16, "Ptr": 8}
C_WIDTH = {"int64_t": 8, "int32_t": 4, "double": 8, "mfd_c64": 16}

# exports of the reference, src/model/model.jl:23-28
REFERENCE_EXPORTS = ["set_ωpml!", "set_boundft!", "set_Npml!", "set_kbloch!", "create_e⁻ⁱᵏᴸ", "clear_objs!",
                     "clear_srcs!", "add_srce!", "add_srcm!", "create_paramops", "create_curls", "create_srcs",
                     "create_linsys", "create_A", "create_b", "h_from_e", "e_from_h", "create_Mcs"]


def source():
    return open(JL, encoding="utf-8").read()


def header():
    return re.sub(r"/\*.*?\*/", "", open(HEADER, encoding="utf-8").read(), flags=re.S)


def julia_struct(name):
    m = re.search(r"struct " + name + r"\b[^\n]*\n(.*?)\nend", source(), flags=re.S)
    assert m, name
    out = []
    for line in m.group(1).splitlines():
        line = line.strip()
        if not line or line.startswith("#") or "::" not in line or line.startswith("function"):
            continue
        field, typ = [t.strip() for t in line.split("::", 1)]
        typ = typ.split("#")[0].strip()
        tm = re.match(r"NTuple\{(\d+),\s*NTuple\{(\d+),\s*(\w+)", typ)
        if tm:
            out.append((field, int(tm.group(1)) * int(tm.group(2)) * JULIA_WIDTH[tm.group(3)]))
            continue
        tm = re.match(r"NTuple\{(\d+),\s*(\w+)", typ)
        if tm:
            out.append((field, int(tm.group(1)) * JULIA_WIDTH[tm.group(2)]))
            continue
        out.append((field, JULIA_WIDTH[re.match(r"(\w+)", typ).group(1)]))
    return out


def c_struct(name):
    m = re.search(r"typedef struct \{([^{}]*)\}\s*" + name + ";", header())
    assert m, name
    out = []
    for decl in m.group(1).split(";"):
        decl = decl.strip()
        if not decl:
            continue
        typ = decl.split()[0] if not decl.startswith("const") else decl.split()[1]
        ptr = "*" in decl
        for part in decl.split(",") if not ptr else [decl]:
            nm = re.findall(r"(\w+)\s*((?:\[\d+\])*)\s*$", part.strip())
            if ptr:
                names = re.findall(r"\*\s*(\w+)((?:\[\d+\])*)", decl)
                for n_, dims in names:
                    k = 1
                    for d in re.findall(r"\d+", dims):
                        k *= int(d)
                    out.append((n_, 8 * k))
                break
            n_, dims = nm[0]
            k = 1
            for d in re.findall(r"\d+", dims):
                k *= int(d)
            out.append((n_, C_WIDTH[typ] * k))
    return out


def test_ccall_symbols_declared_and_exported(pkg):
    syms = sorted(set(re.findall(r"ccall\(\(:(mfd_\w+),\s*libmfdfd\)", source())))
    assert len(syms) >= 15
    declared = set(re.findall(r"\b(mfd_\w+)\s*\(", header()))
    raw = C.CDLL(pkg.LIB_PATH)
    for s in syms:
        assert s in declared, f"{s} is not declared in include/mfdfd.h"
        assert hasattr(raw, s), f"{s} is not exported by libmfdfd"


def test_struct_layouts_match_header():
    for jl, c in (("MfdGrid", "mfd_grid"), ("MfdCsc", "mfd_csc"), ("MfdSolveOpts", "mfd_solve_opts"),
                  ("MfdSolveStats", "mfd_solve_stats")):
        j = julia_struct(jl)
        h = c_struct(c)
        assert [w for _, w in j] == [w for _, w in h], (jl, j, h)


def test_reference_exports_are_exported():
    src = source()
    exported = set()
    for m in re.finditer(r"^export (.*?)(?=\n(?:export|\S))", src, flags=re.S | re.M):
        exported |= {t.strip() for t in m.group(1).replace("\n", " ").split(",") if t.strip()}
    missing = [n for n in REFERENCE_EXPORTS if n not in exported]
    assert not missing, missing
    for n in ("solve_linsys", "solve_field", "MaxwellOperator"):
        assert n in exported
    assert "ChainRulesCore.rrule(::typeof(solve_field)" in src
