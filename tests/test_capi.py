"""The C-ABI library loads and exports every symbol include/vortexhip.h declares (no GPU needed)."""
import ctypes as C
import os
import re

from vphip import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "vortexhip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char \*)\s*(vp_\w+)\(", src, re.M)))


def test_header_matches_binding_list():
    assert declared() == sorted(_lib.EXPORTS)


def test_library_exports_every_symbol():
    lib = C.CDLL(_lib.LIBPATH)
    for name in declared():
        assert hasattr(lib, name), name


def _c_structs():
    """Field names and types of every `typedef struct vp_X { ... } vp_X;` in include/vortexhip.h"""
    import re
    txt = open(os.path.join(ROOT, "include", "vortexhip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    out = {}
    for name, body in re.findall(r"typedef struct (vp_\w+) \{(.*?)\} \1;", txt, flags=re.S):
        fields = []
        for decl in body.split(";"):
            decl = " ".join(decl.split())
            if not decl:
                continue
            m = re.match(r"(const )?((?:struct )?(?:long long|unsigned \w+|\w+))\s+(.*)$", decl)
            assert m, (name, decl)
            for d in m.group(3).split(","):
                dm = re.match(r"\s*(\*?)\s*(\w+)(\[(\d+)\])?\s*$", d)
                assert dm, (name, decl)
                fields.append((dm.group(2), m.group(2).strip(), bool(dm.group(1)), dm.group(4)))
        out[name] = fields
    return out


def test_julia_structs_mirror_the_header():
    """The Julia shim's C structs (VortexPanelHIP.jl) list the header's fields in order with matching types"""
    import re
    jl = open(os.path.join(ROOT, "vortexpanel.jl_amd", "julia", "VortexPanelHIP.jl")).read()
    cmap = {"int": "Cint", "double": "Cdouble", "long long": "Clonglong", "uint8_t": "UInt8"}
    structs = _c_structs()
    found = 0
    for jname, cname, body in re.findall(r"\nstruct (VP\w+)\s+# (vp_\w+)\n(.*?)\nend", jl, flags=re.S):
        found += 1
        jf = [tuple(l.strip().split("::")) for l in body.strip().splitlines()]
        cf = structs[cname]
        assert [f[0] for f in jf] == [f[0] for f in cf], (jname, jf, cf)
        for (fname, jt), (_, ct, ptr, arr) in zip(jf, cf):
            want = cmap.get(ct, "Cvoid")
            if ptr:
                assert jt.startswith("Ptr{"), (jname, fname, jt)
                if ct in cmap:
                    assert jt == f"Ptr{{{want}}}", (jname, fname, jt)
            elif arr:
                assert jt == f"NTuple{{{arr},{want}}}", (jname, fname, jt)
            else:
                assert jt == want, (jname, fname, jt, ct)
    assert found >= 5
