"""The Julia binding (testparticle.jl_amd/julia) against the C header, without
Julia: the struct mirrors must lay out like the C structs (offsets from the C
compiler), every ccall must name an exported entry point with the header's
parameter count and kinds, and the algorithm mapping must follow the types,
not the safety value (MultistepBoris{N, false}(; n) carries safety = 0.1,
src/types.jl:70-76)."""
import ctypes as C
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
JL = os.path.join(ROOT, "testparticle.jl_amd", "julia")
SRC = os.path.join(JL, "src", "TestParticleAMD.jl")
EXT = os.path.join(JL, "ext", "TestParticleAMDChainRulesCoreExt.jl")
HDR = os.path.join(ROOT, "include", "tpamd.h")

JL_SIZES = {"Int32": (4, 4), "Int64": (8, 8), "Float64": (8, 8), "NTuple{3, Float64}": (24, 8)}


def _jl_struct(src, name):
    m = re.search(r"(?:Base\.@kwdef\s+)?mutable struct " + name + r"\n(.*?)\nend", src, re.S)
    assert m, name
    fields = []
    for line in m.group(1).splitlines():
        line = line.split("#")[0].strip()
        if not line:
            continue
        f, t = line.split("::")
        t = t.split("=")[0].strip()
        fields.append((f.strip(), "Ptr" if t.startswith("Ptr{") else t))
    return fields


def _c_layout(fields):
    off, out, align_max = 0, [], 1
    for f, t in fields:
        size, al = (8, 8) if t == "Ptr" else JL_SIZES[t]
        off = (off + al - 1) // al * al
        out.append((f, off))
        off += size
        align_max = max(align_max, al)
    return out, (off + align_max - 1) // align_max * align_max


def _have_gcc():
    return any(os.access(os.path.join(d, "gcc"), os.X_OK) for d in os.environ.get("PATH", "").split(os.pathsep))


@pytest.mark.skipif(not _have_gcc(), reason="no C compiler")
def test_julia_struct_mirrors_match_the_c_header(tmp_path):
    src = open(SRC).read()
    structs = {"tpamd_solve_desc": _jl_struct(src, "SolveDesc"), "tpamd_stats": _jl_struct(src, "Stats")}
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "tpamd.h"', "int main(void) {"]
    for cname, fields in structs.items():
        lines.append(f'printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for f, _t in fields:
            lines.append(f'printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["return 0;", "}"]
    (tmp_path / "l.c").write_text("\n".join(lines))
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", str(tmp_path / "l"),
                    str(tmp_path / "l.c")], check=True)
    out = subprocess.run([str(tmp_path / "l")], capture_output=True, text=True, check=True).stdout
    got = {(c, n): int(v) for c, n, v in (l.split() for l in out.splitlines() if l)}
    for cname, fields in structs.items():
        offs, size = _c_layout(fields)
        assert got[(cname, "sizeof")] == size, cname
        for f, o in offs:
            assert got[(cname, f)] == o, (cname, f)
    # every C field is mirrored (no missing member such as stats.aux_ms)
    hdr = open(HDR).read()
    for cname, jname in (("tpamd_solve_desc", "SolveDesc"), ("tpamd_stats", "Stats")):
        body = re.search(r"typedef struct " + cname + r" \{(.*?)\} " + cname, hdr, re.S).group(1)
        cfields = []
        for decl in re.sub(r"/\*.*?\*/", "", body, flags=re.S).split(";"):
            decl = decl.strip()
            if not decl:
                continue
            # "double t0, t1" / "double box_lo[3], box_hi[3]" / "const double* dt_pp"
            first, *rest = [d.strip() for d in decl.split(",")]
            names = [re.search(r"(\w+)(?:\[\d+\])?$", first).group(1)]
            names += [re.match(r"\**(\w+)", d).group(1) for d in rest]
            cfields += names
        assert cfields == [f for f, _ in structs[cname]], cname


def _c_prototypes():
    hdr = re.sub(r"/\*.*?\*/", "", open(HDR).read(), flags=re.S)
    protos = {}
    for m in re.finditer(r"\n(?:const )?[\w]+\**\s*\**(tpamd_\w+)\((.*?)\);", hdr, re.S):
        params = [p.strip() for p in m.group(2).replace("\n", " ").split(",")]
        if params == ["void"]:
            params = []
        kinds = []
        for p in params:
            if "*" in p or "[" in p:
                kinds.append("ptr")
            elif p.startswith("int32_t"):
                kinds.append("Int32")
            elif p.startswith("int64_t"):
                kinds.append("Int64")
            elif p.startswith("double"):
                kinds.append("Float64")
            else:
                raise AssertionError(p)
        protos[m.group(1)] = kinds
    return protos


def _split_top(s):
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch in "({[":
            depth += 1
        elif ch in ")}]":
            depth -= 1
        if ch == "," and depth == 0:
            out.append(cur.strip())
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur.strip())
    return out


def test_julia_ccalls_match_the_c_prototypes():
    protos = _c_prototypes()
    src = open(SRC).read()
    calls = re.findall(r"ccall\(\(:(tpamd_\w+), libtpamd\),\s*(\w+),\s*\((.*?)\),", src, re.S)
    assert len(calls) >= 10
    for name, ret, args in calls:
        assert name in protos, name
        jl = [a for a in _split_top(args.replace("\n", " ")) if a]
        kinds = ["ptr" if a.startswith(("Ptr{", "Ref{", "Cstring")) else a for a in jl]
        assert kinds == protos[name], (name, kinds, protos[name])
        assert ret in ("Int32", "Cstring"), name
    # and the library exports each of them
    from testparticle_amd import _lib  # noqa: F401
    lib = C.CDLL(os.path.join(ROOT, "testparticle.jl_amd", "_build", "libtpamd.so"))
    for name, _r, _a in calls:
        assert hasattr(lib, name), name


def test_julia_algorithm_mapping_follows_the_types():
    src = open(SRC).read()
    assert "alg.safety > 0" not in src
    assert re.search(r"is_adaptive\(::Boris\{A\}\) where \{A\} = A", src)
    assert re.search(r"is_adaptive\(::MultistepBoris\{N, A\}\) where \{N, A\} = A", src)
    assert "d.adaptive = is_adaptive(alg) ? 1 : 0" in src


def test_rrule_extension_is_declared_and_defined():
    proj = open(os.path.join(JL, "Project.toml")).read()
    assert 'ChainRulesCore = "d360d2e6-b24c-11e9-a2a3-2a2ae2dbcce4"' in proj
    assert 'TestParticleAMDChainRulesCoreExt = "ChainRulesCore"' in proj
    ext = open(EXT).read()
    assert "function ChainRulesCore.rrule(::typeof(TestParticleAMD.trajectory_loss)" in ext
    assert "TestParticleAMD.loss_gradient(A, xv0, targets; kw...)" in ext
    src = open(SRC).read()
    assert "function loss_gradient(" in src and "trajectory_loss(A, xv0, targets; kw...)" in src
