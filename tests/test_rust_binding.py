"""The Rust binding crate (crates/syfrack-hip-sys) against the C header, on the CPU.

No Rust toolchain is in this image, so the crate's raw layer is checked by reading it:
  * every `#[repr(C)]` structure of src/lib.rs is rewritten field for field as a C struct and
    compiled by gcc beside include/syfrack_hip.h with static assertions that its size, alignment
    and every field offset equal the header's structure of the same name, and its field names
    are the header's, in order (a dropped trailing field can hide in padding);
  * the extern block declares exactly the header's functions, each with the header's parameter
    count;
  * every integer constant of the header has a Rust constant of the same name and value.
"""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_RS = os.path.join(ROOT, "crates", "syfrack-hip-sys", "src", "lib.rs")
SCENE_RS = os.path.join(ROOT, "crates", "syfrack-hip-sys", "src", "scene.rs")
HEADER = os.path.join(ROOT, "include", "syfrack_hip.h")

SCALAR = {"u8": "uint8_t", "u32": "uint32_t", "i32": "int32_t", "u64": "uint64_t", "f32": "float", "f64": "double",
          "usize": "size_t", "c_int": "int", "c_char": "char"}


def rust_consts(text):
    out = {}
    for name, value in re.findall(r"pub const (\w+): \w+ = ([0-9_xa-fA-F-]+);", text):
        out[name] = int(value.replace("_", ""), 0)
    return out


def rust_structs(text):
    out = {}
    for name, body in re.findall(r"#\[repr\(C\)\]\s*(?:#\[derive\([^)]*\)\]\s*)?pub struct (\w+) \{(.*?)\n\}", text, re.S):
        fields = re.findall(r"pub (\w+): ([^,\n]+),", body)
        if fields:
            out[name] = fields
    return out


def c_decl(name, ty, consts):
    ty = ty.strip()
    m = re.match(r"\[(.+); (\w+)\]$", ty)
    if m:
        n = m.group(2)
        n = consts[n] if n in consts else int(n)
        return c_decl(f"{name}[{n}]", m.group(1), consts)
    m = re.match(r"\*(const|mut) (\w+)$", ty)
    if m:
        base = SCALAR.get(m.group(2), m.group(2))
        return f"{'const ' if m.group(1) == 'const' else ''}{base}* {name}"
    return f"{SCALAR.get(ty, ty)} {name}"


def test_repr_c_structs_match_the_header(tmp_path):
    text = open(LIB_RS, encoding="utf-8").read()
    consts = rust_consts(text)
    structs = rust_structs(text)
    assert {"syf_fractal", "syf_camera", "syf_light", "syf_material", "syf_graph_node", "syf_graph_link", "syf_graph",
            "syf_job_desc", "syf_stats", "syf_planar_desc", "syf_server_config", "syf_request_info",
            "syf_server_counters"} <= set(structs)
    lines = ["#include <stddef.h>", "#include <stdint.h>", f'#include "{HEADER}"']
    for name, fields in structs.items():
        lines.append("typedef struct {")
        for f, ty in fields:
            lines.append("    " + c_decl(f, ty, consts) + ";")
        lines.append(f"}} rs_{name};")
        lines.append(f'_Static_assert(sizeof(rs_{name}) == sizeof({name}), "size of {name}");')
        lines.append(f'_Static_assert(_Alignof(rs_{name}) == _Alignof({name}), "alignment of {name}");')
        for f, _ in fields:
            lines.append(f'_Static_assert(offsetof(rs_{name}, {f}) == offsetof({name}, {f}), "{name}.{f}");')
    c = tmp_path / "layout.c"
    c.write_text("\n".join(lines) + "\n")
    r = subprocess.run(["gcc", "-std=c11", "-fsyntax-only", "-Wall", str(c)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[:3000]


def header_struct_fields():
    """Field names of every `typedef struct NAME { ... } NAME;` of the header, in order."""
    text = re.sub(r"/\*.*?\*/", "", open(HEADER, encoding="utf-8").read(), flags=re.S)
    out = {}
    for name, body in re.findall(r"typedef struct (\w+) \{(.*?)\} \1;", text, re.S):
        fields = []
        for stmt in body.split(";"):
            stmt = stmt.strip()
            if not stmt:
                continue
            m = re.match(r"(?:const\s+)?\w+\s*\*?\s*(.+)$", stmt, re.S)
            for decl in m.group(1).split(","):
                fields.append(re.sub(r"\[.*?\]", "", decl).strip().lstrip("*").strip())
        out[name] = fields
    return out


def test_repr_c_structs_have_the_header_fields_in_order():
    structs = rust_structs(open(LIB_RS, encoding="utf-8").read())
    header = header_struct_fields()
    assert set(structs) == set(header), (sorted(set(header) - set(structs)), sorted(set(structs) - set(header)))
    for name, fields in structs.items():
        assert [f for f, _ in fields] == header[name], name


def header_functions():
    text = open(HEADER, encoding="utf-8").read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    out = {}
    for m in re.finditer(r"^\s*(?:const\s+)?\w+\s*\*?\s*(syf_\w+)\s*\(([^;]*?)\);", text, re.M | re.S):
        params = m.group(2).strip()
        out[m.group(1)] = 0 if params in ("", "void") else params.count(",") + 1
    return out


def rust_functions(text):
    block = text[text.index('extern "C" {'):]
    out = {}
    for m in re.finditer(r"pub fn (syf_\w+)\((.*?)\)\s*(?:->[^;]*)?;", block, re.S):
        params = m.group(2).strip().rstrip(",")
        out[m.group(1)] = 0 if not params else params.count(",") + 1
    return out


def test_extern_block_declares_the_header_functions():
    text = open(LIB_RS, encoding="utf-8").read()
    h, r = header_functions(), rust_functions(text)
    assert set(h) == set(r), (sorted(set(h) - set(r)), sorted(set(r) - set(h)))
    for name in h:
        assert h[name] == r[name], f"{name}: {h[name]} parameters in the header, {r[name]} in lib.rs"


def test_constants_match_the_header():
    text = open(LIB_RS, encoding="utf-8").read()
    consts = rust_consts(text)
    header = open(HEADER, encoding="utf-8").read()
    defines = {}
    for name, value in re.findall(r"^#define (SYF_\w+)\s+\(?(-?(?:0x[0-9a-fA-F]+|\d+))u?\)?", header, re.M):
        defines[name] = int(value, 0)
    missing = sorted(set(defines) - set(consts))
    assert not missing, missing
    for name, value in defines.items():
        assert consts[name] == value, name


def test_scene_lowering_names_every_node_kind():
    """The safe layer lowers every NodeKind variant (mod.rs:103-257) to its SYF_NODE_* number."""
    text = open(SCENE_RS, encoding="utf-8").read()
    header = open(HEADER, encoding="utf-8").read()
    kinds = re.findall(r"#define (SYF_NODE_\w+) \d+u", header)
    kinds = [k for k in kinds if k != "SYF_NODE_KINDS"]
    assert len(kinds) == 41
    for k in kinds:
        assert re.search(rf"\b{k}\b", text), k
