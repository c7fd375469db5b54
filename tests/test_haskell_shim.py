"""The Haskell binding (lulesh-accelerate_amd/haskell/LULESH/MI355X.hs)
against the C ABI it binds (include/lulesh_mi355x.h).  GHC is not available
here, so what the module hard-codes is checked from its source:

- every `foreign import ccall` names a function the header declares, with
  the same argument and result types (CInt = int, CDouble = double,
  CString = char *, Ptr T = T *, FunPtr for the finalizer);
- the Storable instances of Config and Stats: size, alignment, and the
  byte offset and C type of every record field, against sizeof, _Alignof and
  offsetof of the header's structs (compiled with gcc);
- the return codes and the field identifiers, against the header's values.

The GPU test then drives the library through ctypes prototypes built from
the module's own foreign imports and struct offsets, in the module's call
sequence (config_default, create, step, get_stats, download_state,
destroy), and compares the result with the Python binding's run."""
import ctypes
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lulesh_mi355x.h")
MODULE = os.path.join(ROOT, "lulesh-accelerate_amd", "haskell", "LULESH", "MI355X.hs")

STRUCTS = {"Config": ("lulesh_config", "cfg_"), "Stats": ("lulesh_stats", "st_")}
CODES = {"codeOk": "LULESH_OK", "codeVolumeError": "LULESH_VOLUME_ERROR",
         "codeQStopError": "LULESH_QSTOP_ERROR", "codeBadArg": "LULESH_BAD_ARG",
         "codeHipError": "LULESH_HIP_ERROR"}
HS_SCALARS = {"CInt": "int", "CDouble": "double", "CFloat": "float", "CString": "char*"}
HS_STRUCTS = {"Config": "lulesh_config", "Stats": "lulesh_stats",
              "LuleshDomain": "lulesh_domain"}


def hs_source() -> str:
    text = open(MODULE).read()
    text = re.sub(r"\{-.*?-\}", "", text, flags=re.S)
    return "\n".join(line.split("--", 1)[0].rstrip() for line in text.splitlines())


def split_arrows(t: str) -> list[str]:
    """Splits a Haskell function type at its top-level arrows."""
    parts, depth, cur = [], 0, ""
    i = 0
    while i < len(t):
        c = t[i]
        if c == "(":
            depth += 1
        elif c == ")":
            depth -= 1
        if depth == 0 and t.startswith("->", i):
            parts.append(cur.strip())
            cur = ""
            i += 2
            continue
        cur += c
        i += 1
    parts.append(cur.strip())
    return parts


def hs_to_c(t: str) -> str:
    t = " ".join(t.split())
    while t.startswith("(") and t.endswith(")"):
        t = t[1:-1].strip()
    if t == "IO ()":
        return "void"
    if t.startswith("IO "):
        return hs_to_c(t[3:])
    if t in HS_SCALARS:
        return HS_SCALARS[t]
    if t.startswith("Ptr "):
        inner = t[4:].strip()
        while inner.startswith("(") and inner.endswith(")"):
            inner = inner[1:-1].strip()
        if inner in HS_STRUCTS:
            return HS_STRUCTS[inner] + "*"
        return hs_to_c(inner) + "*"
    raise ValueError(f"no C type for Haskell type {t!r}")


def foreign_imports() -> list[dict]:
    out = []
    src = hs_source()
    for m in re.finditer(r'foreign import ccall (?:safe|unsafe) "(&?)(\w+)"\s*\n\s*(\w+)\s*::'
                         r"(.*?)(?=\n\S|\Z)", src, flags=re.S):
        addr, cname, hsname, ty = m.group(1), m.group(2), m.group(3), m.group(4)
        ty = " ".join(ty.split())
        if addr:
            fm = re.fullmatch(r"FunPtr \((.*)\)", ty)
            assert fm, f"{hsname}: an address import must be a FunPtr"
            parts = split_arrows(fm.group(1))
        else:
            parts = split_arrows(ty)
        out.append({"c": cname, "hs": hsname, "addr": bool(addr),
                    "args": [hs_to_c(p) for p in parts[:-1]], "ret": hs_to_c(parts[-1])})
    return out


def header_functions() -> dict:
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    text = re.sub(r"^\s*#.*$", "", text, flags=re.M)
    decls = {}
    for m in re.finditer(r"([A-Za-z_][\w\s\*]*?)\b(lulesh_\w+)\s*\(([^)]*)\)\s*;", text):
        ret = normalise(m.group(1))
        args = [] if m.group(3).strip() in ("", "void") else [
            normalise(re.sub(r"\b\w+\s*$", "", a.strip())) for a in m.group(3).split(",")]
        decls[m.group(2)] = {"ret": ret, "args": args}
    return decls


def normalise(c: str) -> str:
    c = re.sub(r"\bconst\b", "", c)
    c = re.sub(r"\bstruct\b", "", c)
    stars = c.count("*")
    return " ".join(c.replace("*", " ").split()) + "*" * stars


def test_foreign_imports_match_the_header():
    imports = foreign_imports()
    decls = header_functions()
    assert len(imports) >= 12
    for imp in imports:
        assert imp["c"] in decls, f"{imp['c']} is not declared in the header"
        d = decls[imp["c"]]
        assert imp["args"] == d["args"], (imp["c"], imp["args"], d["args"])
        assert imp["ret"] == d["ret"], (imp["c"], imp["ret"], d["ret"])
    # the calls the drop-in for `simulate` rests on
    names = {i["c"] for i in imports}
    for need in ("lulesh_create", "lulesh_step", "lulesh_run", "lulesh_get_stats",
                 "lulesh_download_state", "lulesh_upload_state", "lulesh_destroy"):
        assert need in names, need


def storable(name: str) -> dict:
    """Record fields (in order, with types), sizeOf, alignment and the
    offsets of the Storable instance of `name` in the module."""
    src = hs_source()
    rec = re.search(r"data %s = %s\s*\{(.*?)\}" % (name, name), src, flags=re.S).group(1)
    fields = re.findall(r"(\w+)\s*::\s*!?(\w+)", rec)
    inst = re.search(r"instance Storable %s where(.*?)(?=\n\S|\Z)" % name, src,
                     flags=re.S).group(1)
    size = int(re.search(r"sizeOf _\s*=\s*(\d+)", inst).group(1))
    align = int(re.search(r"alignment _\s*=\s*(\d+)", inst).group(1))
    peek_offs = [int(x) for x in re.findall(r"peekByteOff p (\d+)", inst)]
    poke = dict((f, int(o)) for o, f in re.findall(r"pokeByteOff p (\d+)\s*\((\w+) \w+\)",
                                                   inst))
    return {"fields": fields, "size": size, "align": align, "peek": peek_offs, "poke": poke}


def c_facts(tmp_path) -> dict:
    """sizeof, _Alignof, offsetof and the field types of the header's
    structs, and the values of its return codes and field identifiers."""
    text = open(HEADER).read()
    enum_names = sorted(set(re.findall(r"\b(LULESH_F_\w+)\b", text)))
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "lulesh_mi355x.h"',
             "#define T(x) _Generic((x), int: \"int\", double: \"double\", default: \"?\")",
             "int main(void) {"]
    for hs, (cname, prefix) in STRUCTS.items():
        lines.append(f'printf("size {cname} %zu %zu\\n", sizeof({cname}), _Alignof({cname}));')
        for field, _ in storable(hs)["fields"]:
            cf = field[len(prefix):]
            lines.append(f'{{ {cname} v; printf("field {cname} {cf} %zu %s\\n", '
                         f'offsetof({cname}, {cf}), T(v.{cf})); }}')
    for code in CODES.values():
        lines.append(f'printf("code {code} %d\\n", {code});')
    for e in enum_names:
        lines.append(f'printf("enum {e} %d\\n", (int){e});')
    lines.append("return 0; }")
    src = tmp_path / "facts.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "facts"
    subprocess.run(["gcc", "-std=c11", "-I" + os.path.dirname(HEADER), str(src), "-o", str(exe)],
                   check=True, capture_output=True, text=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    facts = {"size": {}, "field": {}, "code": {}, "enum": {}}
    for line in out.splitlines():
        k = line.split()
        if k[0] == "size":
            facts["size"][k[1]] = (int(k[2]), int(k[3]))
        elif k[0] == "field":
            facts["field"][(k[1], k[2])] = (int(k[3]), k[4])
        elif k[0] == "code":
            facts["code"][k[1]] = int(k[2])
        else:
            facts["enum"][k[1]] = int(k[2])
    return facts


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_struct_layouts_match_the_header(tmp_path):
    facts = c_facts(tmp_path)
    hs_c = {"CInt": "int", "CDouble": "double"}
    for hs, (cname, prefix) in STRUCTS.items():
        st = storable(hs)
        assert (st["size"], st["align"]) == facts["size"][cname], (hs, facts["size"][cname])
        # every C field has a Haskell field and no Haskell field is extra
        c_fields = {f for (s, f) in facts["field"] if s == cname}
        assert {f[len(prefix):] for f, _ in st["fields"]} == c_fields
        assert len(st["peek"]) == len(st["fields"])
        for (field, ty), peek_off in zip(st["fields"], st["peek"]):
            off, cty = facts["field"][(cname, field[len(prefix):])]
            assert peek_off == off, (hs, field, peek_off, off)
            assert st["poke"][field] == off, (hs, field, st["poke"][field], off)
            assert hs_c[ty] == cty, (hs, field, ty, cty)
    # a struct that grows or reorders fields is caught by the checks above;
    # check the header's own count as well
    assert len(STRUCTS) == 2


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_codes_and_field_ids_match_the_header(tmp_path):
    facts = c_facts(tmp_path)
    src = hs_source()
    for hs, c in CODES.items():
        m = re.search(r"^%s\s*=\s*(-?\d+)\s*$" % hs, src, flags=re.M)
        assert m, hs
        assert int(m.group(1)) == facts["code"][c], (hs, c)
    ids = dict((n, int(v)) for n, v in re.findall(r"^\s+(F_\w+)\s*->\s*(\d+)\s*$", src,
                                                   flags=re.M))
    assert {"LULESH_" + n for n in ids} == set(facts["enum"]), "field constructors"
    for n, v in ids.items():
        assert facts["enum"]["LULESH_" + n] == v, n


def ctypes_from_module(lib):
    """Prototypes of the module's foreign imports (C types from its Haskell
    types) on the library."""
    ct = {"int": ctypes.c_int, "double": ctypes.c_double, "void": None,
          "char*": ctypes.c_char_p}
    protos = {}
    for imp in foreign_imports():
        fn = getattr(lib, imp["c"])
        fn.argtypes = [ct.get(a, ctypes.c_void_p) if not a.endswith("*") or a == "char*"
                       else ctypes.c_void_p for a in imp["args"]]
        fn.restype = ct.get(imp["ret"], ctypes.c_void_p)
        protos[imp["c"]] = fn
    return protos


@pytest.mark.gpu
def test_module_call_sequence_on_the_gpu(lulesh_gpu):
    """newDomain, step, stats, downloadState and the finalizer, as the module
    makes them, with the module's struct offsets."""
    p = ctypes_from_module(ctypes.CDLL(lulesh_gpu.LIB_PATH))
    cfg_l = storable("Config")
    st_l = storable("Stats")
    cfg = (ctypes.c_char * cfg_l["size"])()
    p["lulesh_config_default"](ctypes.addressof(cfg))
    put = dict(cfg_l["poke"])
    ctypes.c_int.from_buffer(cfg, put["cfg_nx"]).value = 12
    ctypes.c_int.from_buffer(cfg, put["cfg_ic_mode"]).value = 1
    ctypes.c_int.from_buffer(cfg, put["cfg_max_cycles"]).value = 40
    h = ctypes.c_void_p()
    assert p["lulesh_create"](ctypes.addressof(cfg), ctypes.addressof(h)) == 0
    n = ctypes.c_int()
    assert p["lulesh_step"](h, 25, ctypes.addressof(n)) == 0 and n.value == 25
    st = (ctypes.c_char * st_l["size"])()
    assert p["lulesh_get_stats"](h, ctypes.addressof(st)) == 0
    get = dict(st_l["poke"])
    cycle = ctypes.c_int.from_buffer(st, get["st_cycle"]).value
    num_node = ctypes.c_int.from_buffer(st, get["st_num_node"]).value
    num_elem = ctypes.c_int.from_buffer(st, get["st_num_elem"]).value
    time_ = ctypes.c_double.from_buffer(st, get["st_time"]).value
    assert (cycle, num_node, num_elem) == (25, 13 ** 3, 12 ** 3)
    arrs = [np.empty(num_node) for _ in range(6)] + [np.empty(num_elem) for _ in range(5)]
    t, dt = ctypes.c_double(), ctypes.c_double()
    assert p["lulesh_download_state"](h, *[a.ctypes.data for a in arrs],
                                      ctypes.addressof(t), ctypes.addressof(dt)) == 0
    assert t.value == time_
    # the finalizer the module registers
    destroy = ctypes.CFUNCTYPE(None, ctypes.c_void_p)(
        ctypes.cast(ctypes.CDLL(lulesh_gpu.LIB_PATH).lulesh_destroy, ctypes.c_void_p).value)
    destroy(h)
    # the same run through the Python binding
    d = lulesh_gpu.Domain(12, ic_mode=1, max_cycles=40)
    assert d.step(25) == (0, 25)
    ref = d.download_state()
    for a, k in zip(arrs, ("x", "y", "z", "xd", "yd", "zd", "e", "p", "q", "v", "ss")):
        assert np.array_equal(a.view(np.uint64), ref[k].view(np.uint64)), k
    d.close()
