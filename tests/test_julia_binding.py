"""The Julia binding a maintainer adds (the `HesGPU` module of the
integration notes at the repository root, NOTES below) against the C-ABI it
calls: every `ccall` names a function include/hes.h
declares, passes as many arguments as it lists types, and lists the types of
the C prototype in order (Julia is absent from this image, so this is the
check that the stub would link and pass arguments correctly).  The module
also defines every function of the reference's module API that the library
replaces."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NOTES = os.path.join(ROOT, "INTEGRATION.md")


def _strip_c_comments(text):
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.sub(r"//[^\n]*", "", text)


def c_prototypes():
    """name -> (return type, [parameter types]) of include/hes.h."""
    text = _strip_c_comments(open(os.path.join(ROOT, "include", "hes.h")).read())
    out = {}
    for m in re.finditer(r"([A-Za-z_][A-Za-z0-9_ \*]*?)\b(hes_[a-z0-9_]+)\s*\(([^)]*)\)\s*;", text):
        ret = " ".join(m.group(1).split())
        params = [p.strip() for p in m.group(3).split(",")] if m.group(3).strip() not in ("", "void") else []
        types = []
        for p in params:
            p = " ".join(p.replace("*", " * ").split())
            toks = p.split(" ")
            # drop the parameter name (last identifier)
            if toks and re.fullmatch(r"[A-Za-z_][A-Za-z0-9_]*", toks[-1]) and len(toks) > 1:
                toks = toks[:-1]
            types.append(" ".join(toks))
        out[m.group(2)] = (ret, types)
    return out


def _balanced(text, start):
    """Index just past the parenthesis that closes the one at text[start]."""
    depth = 0
    for k in range(start, len(text)):
        if text[k] == "(":
            depth += 1
        elif text[k] == ")":
            depth -= 1
            if depth == 0:
                return k + 1
    raise ValueError("unbalanced")


def _split_top(text):
    """Split on commas outside (), {} and []."""
    parts, depth, cur = [], 0, ""
    for ch in text:
        if ch in "({[":
            depth += 1
        elif ch in ")}]":
            depth -= 1
        if ch == "," and depth == 0:
            parts.append(cur.strip())
            cur = ""
        else:
            cur += ch
    if cur.strip():
        parts.append(cur.strip())
    return parts


def julia_block():
    text = open(NOTES).read()
    m = re.search(r"```julia\n(module HesGPU.*?)```", text, re.S)
    assert m, "no HesGPU module in the integration notes"
    return m.group(1)


def ccalls(code):
    """[(name, return type, [argument types], [arguments])] of every ccall."""
    out = []
    for m in re.finditer(r"ccall\(", code):
        end = _balanced(code, m.end() - 1)
        parts = _split_top(code[m.end():end - 1])
        name = re.fullmatch(r"\(:(\w+),\s*(?:HesGPU\.)?LIB\)", parts[0]).group(1)
        ret = parts[1]
        tup = parts[2]
        assert tup.startswith("(") and tup.endswith(")"), tup
        types = [t for t in _split_top(tup[1:-1]) if t]
        out.append((name, ret, types, parts[3:]))
    return out


# Julia argument type -> the C parameter types it may stand for
_JULIA_TO_C = {
    "Cint": {"int"},
    "Clonglong": {"long long"},
    "Cdouble": {"double"},
    "Float64": {"double"},
    "Cfloat": {"float"},
    "Ptr{Float64}": {"double *", "const double *"},
    "Ref{Float64}": {"double *", "const double *"},
    "Ref{Cdouble}": {"double *", "const double *"},
    "Ref{Clonglong}": {"long long *", "const long long *"},
    "Ptr{Clonglong}": {"long long *", "const long long *"},
    "Ptr{Cint}": {"int *", "const int *"},
    "Ref{Cint}": {"int *", "const int *"},
    "Ptr{Cdouble}": {"double *", "const double *"},
    "Ref{Cfloat}": {"float *"},
    "Ref{Ptr{Cvoid}}": {"hes_solver * *", "hes_batch * *", "hes_adjoint * *", "hes_badjoint * *"},
    "Ptr{Cvoid}": {"hes_solver *", "hes_batch *", "hes_adjoint *", "hes_badjoint *", "void *",
                   "const void *"},
}
_RET = {"Cint": "int", "Cstring": "const char *"}


def all_julia_blocks():
    text = open(NOTES).read()
    return "\n".join(re.findall(r"```julia\n(.*?)```", text, re.S))


def test_every_ccall_matches_the_c_prototype():
    protos = c_prototypes()
    calls = ccalls(all_julia_blocks())  # the module and the batched / adjoint bindings
    assert len(calls) >= 40
    for name, ret, types, args in calls:
        assert name in protos, f"{name} is not declared in include/hes.h"
        cret, ctypes_ = protos[name]
        assert _RET[ret] == " ".join(cret.replace("*", " * ").split()), (name, ret, cret)
        assert len(types) == len(ctypes_), (name, types, ctypes_)
        assert len(args) == len(types), (name, "arguments", args, types)
        for k, (jt, ct) in enumerate(zip(types, ctypes_)):
            assert jt in _JULIA_TO_C, (name, jt)
            assert ct in _JULIA_TO_C[jt], (name, k, jt, ct)


def test_header_parser_sees_every_declaration():
    """The prototype parser above reads every declaration (so the check
    cannot pass by skipping one)."""
    from test_cpu_abi import declared_symbols
    assert set(c_prototypes()) == set(declared_symbols())


# the reference functions the binding replaces (module, name) -> Julia method
# names defined in the HesGPU module (Hyperelasticity's in HesGPU.Single)
REPLACED = {
    "HyperelasticityMPh": ["prim2cons_mph", "cons2prim_mph", "flux_mph", "noncons_flux",
                           "get_eigvals", "cons2prim", "flux"],
    "NumFluxes": ["lxf", "hll"],
    "EquationsOfState": ["energy", "entropy", "stress", "acoustic"],
    "Strains": ["finger", "invariants"],
    "Hyperelasticity": ["Single.cons2prim", "Single.prim2cons", "Single.flux"],
    "main.jl": ["run!"],
}


def test_the_module_defines_every_replaced_function():
    code = julia_block()
    single = code[code.index("module Single"):code.index("end # module Single")]
    outer = code.replace(single, "")
    defined = set(re.findall(r"^function ([\w!]+)\(", outer, re.M))
    defined |= {"Single." + n for n in re.findall(r"^function ([\w!]+)\(", single, re.M)}
    for mod, names in REPLACED.items():
        for n in names:
            assert n in defined, (mod, n)
    # two methods each: stress (ent / den, e_int), acoustic (ent / P), get_eigvals (pair / phase)
    for n in ("stress", "acoustic", "get_eigvals"):
        assert len(re.findall(rf"^function {n}\(", outer, re.M)) == 2, n


@pytest.mark.parametrize("name,count", [("hes_sp_cons2prim", 4), ("hes_phase_get_eigvals", 5),
                                        ("hes_solver_run", 8)])
def test_known_prototypes(name, count):
    assert len(c_prototypes()[name][1]) == count
