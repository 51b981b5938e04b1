"""The Rust binding INTEGRATION.md gives for the reference's side (its `extern "C"` block, `#[repr(C)]`
structs and constants) agrees with the C ABI of include/pbrt_gpu.h:

  * every function the listing declares exists in the header with as many parameters, each of the
    same kind (pointer or scalar, and for scalars the same width and signedness / float-ness), and
    both return int;
  * every struct has the same flattened field layout (the sequence of scalar types, arrays
    expanded), so that #[repr(C)] on the Rust side lays it out as the C compiler does;
  * every constant has the header's value.
The listing cannot be compiled here (no Rust toolchain); this is the check that it is the header's.
"""
import os
import re

from conftest import ROOT

C_SCALARS = {"int": "i32", "int32_t": "i32", "uint32_t": "u32", "uint64_t": "u64", "int64_t": "i64", "double": "f64",
             "float": "f32", "size_t": "u64", "uint8_t": "u8", "uint16_t": "u16"}
RUST_SCALARS = {"c_int": "i32", "i32": "i32", "u32": "u32", "u64": "u64", "i64": "i64", "f64": "f64", "f32": "f32",
                "usize": "u64", "u8": "u8", "u16": "u16"}


def _header():
    text = open(os.path.join(ROOT, "include", "pbrt_gpu.h")).read()
    return re.sub(r"/\*.*?\*/", "", text, flags=re.S)


def _rust():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = re.findall(r"```rust\n(.*?)```", text, flags=re.S)
    return re.sub(r"//[^\n]*", "", "\n".join(blocks))


def _c_kind(param):
    param = param.strip()
    if "*" in param or "[" in param:  # an array parameter is a pointer
        return "ptr"
    words = param.replace("const ", "").split()
    return C_SCALARS[words[0]]


def _rust_kind(param):
    ty = param.split(":", 1)[1].strip()
    if ty.startswith("*"):
        return "ptr"
    return RUST_SCALARS[ty]


def header_functions():
    out = {}
    for m in re.finditer(r"(?:int|const char\*)\s+(pbrt_\w+)\s*\(([^;]*?)\)\s*;", _header(), flags=re.S):
        args = m.group(2).strip()
        out[m.group(1)] = [] if args in ("", "void") else [_c_kind(a) for a in args.split(",")]
    return out


def rust_functions():
    out = {}
    for m in re.finditer(r"pub fn (pbrt_\w+)\s*\(([^)]*)\)\s*->\s*([^;]+);", _rust(), flags=re.S):
        args = [a for a in m.group(2).split(",") if a.strip()]
        out[m.group(1)] = ([_rust_kind(a) for a in args], m.group(3).strip())
    return out


def _flatten_c(body):
    seq = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        words = decl.replace("const ", "").replace("*", " * ").replace(",", " , ").split()
        ty = "ptr" if "*" in words else C_SCALARS[words[0]]
        for name in " ".join(w for w in words[1:] if w != "*").split(","):
            name = name.strip()
            n = int(re.search(r"\[(\d+)\]", name).group(1)) if "[" in name else 1
            seq += [ty] * n
    return seq


def _flatten_rust(body):
    seq = []
    for field in body.split(","):
        field = field.strip()
        if not field:
            continue
        ty = field.split(":", 1)[1].strip()
        m = re.match(r"\[(\w+);\s*(\d+)\]", ty)
        seq += [RUST_SCALARS[m.group(1)]] * int(m.group(2)) if m else ["ptr" if ty.startswith("*") else RUST_SCALARS[ty]]
    return seq


def test_listed_functions_match_the_header():
    hdr, rust = header_functions(), rust_functions()
    assert len(rust) >= 30
    for name, (kinds, ret) in rust.items():
        assert name in hdr, f"{name} is listed but not declared"
        if name != "pbrt_last_error":
            assert ret == "c_int", name
        assert kinds == hdr[name], f"{name}: listing {kinds}, header {hdr[name]}"


def test_listed_structs_match_the_header():
    hdr = {m.group(2): _flatten_c(m.group(1))
           for m in re.finditer(r"typedef struct \w+ \{(.*?)\} (\w+);", _header(), flags=re.S)}
    rust = {m.group(1): _flatten_rust(m.group(2))
            for m in re.finditer(r"pub struct (pbrt_\w+)\s*\{([^}]*)\}", _rust(), flags=re.S) if m.group(2).strip() != "_p: [u8; 0]"}
    assert {"pbrt_ray", "pbrt_hit", "pbrt_render_config", "pbrt_launch", "pbrt_scene_hit"} <= set(rust)
    for name, seq in rust.items():
        assert name in hdr, f"{name} is listed but not declared"
        assert seq == hdr[name], f"{name}: listing {seq}, header {hdr[name]}"


def test_listed_constants_match_the_header():
    hdr = dict(re.findall(r"#define (PBRT_\w+) \(?(-?\d+)\)?", _header()))
    consts = re.findall(r"pub const (PBRT_\w+):\s*\w+\s*=\s*(-?\d+);", _rust())
    assert len(consts) >= 10
    for name, value in consts:
        assert name in hdr and int(hdr[name]) == int(value), name
