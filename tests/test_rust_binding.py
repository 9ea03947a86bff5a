"""The Rust binding crate (coral_amd/rust/coral-hip), checked without a Rust toolchain.

There is no cargo or rustc in this image, so the crate is checked by reading it:

* every host-pointer entry point of include/coral_hip.h is declared in ffi.rs with
  the same parameters in the same order and the Rust type of each C type;
* every public function of coral_aarch64's level 1/2/3 modules (the names below, from
  coral-aarch64/src/level1/mod.rs, level2/mod.rs, level3/mod.rs) is defined in the
  module of the same name in lib.rs and bound to the entry point of the same name;
* the enums keep CORAL's variant order (coral-aarch64/src/enums.rs:1-18), which is
  what makes `op as i32` the C ABI's code.
"""
import os
import re

from coral_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CRATE = os.path.join(ROOT, "coral_amd", "rust", "coral-hip")

LEVEL1 = """sasum dasum scasum dzasum isamax idamax icamax izamax snrm2 dnrm2 scnrm2 dznrm2
sscal dscal cscal zscal csscal zdscal scopy dcopy ccopy zcopy sswap dswap cswap zswap
saxpy daxpy caxpy zaxpy sdot ddot cdotu zdotu cdotc zdotc srot srotg srotm srotmg
drot drotg drotm drotmg csrot zdrot""".split()
LEVEL2 = """sgemv dgemv cgemv zgemv strmv dtrmv ctrmv ztrmv strsv dtrsv ctrsv ztrsv ssymv dsymv
chemv zhemv sger dger cgeru cgerc zgeru zgerc ssyr dsyr cher zher ssyr2 dsyr2 cher2 zher2""".split()
LEVEL3 = "sgemm dgemm cgemm zgemm".split()
SPLIT = "sgemm_bf16x3 sgemm_fp16x3 cgemm_3m_bf16x3 cgemm_3m_fp16x3".split()

_RUST = {"int32_t": "i32", "int64_t": "i64", "float": "f32", "double": "f64"}


def _read(*parts):
    with open(os.path.join(CRATE, *parts)) as f:
        return f.read()


def _c_params(params):
    """C parameter list -> [(name, rust type)]."""
    out = []
    for p in params.split(","):
        p = " ".join(p.split())
        m = re.match(r"(const\s+)?(\w+)\s*(\*)?\s*(\w+)(\[\d+\])?$", p)
        assert m, p
        const, base, star, name, arr = m.groups()
        t = _RUST[base]
        if star or arr:
            t = ("*const " if const else "*mut ") + t
        out.append((name, t))
    return out


def _header_host_forms():
    with open(_lib.HEADER_PATH) as f:
        text = re.sub(r"/\*.*?\*/", " ", f.read(), flags=re.S)
    return {m.group(1): _c_params(m.group(2))
            for m in re.finditer(r"\bint\s*(coralh_\w+)\s*\(([^;{]*?)\)\s*;", text, flags=re.S)}


def _ffi_decls():
    text = _read("src", "ffi.rs")
    decls = {}
    for m in re.finditer(r"pub fn (\w+)\((.*?)\)\s*->\s*([\w\s\*]+);", text, flags=re.S):
        params = [tuple(x.strip() for x in p.split(":", 1)) for p in m.group(2).split(",") if p.strip()]
        decls[m.group(1)] = (params, m.group(3).strip())
    return decls


def _module(text, name):
    start = text.index(f"pub mod {name} {{")
    depth = 0
    for i in range(start, len(text)):
        if text[i] == "{":
            depth += 1
        elif text[i] == "}":
            depth -= 1
            if depth == 0:
                return text[start:i + 1]
    raise AssertionError(f"module {name} not closed")


def _bound(module_text):
    """{rust name: ffi name} for the macro invocations `xxx!(name, coralh_name, ...)`."""
    return dict(re.findall(r"(\w+),\s*(coralh_\w+)\s*,", module_text))


def test_ffi_declares_every_host_form_with_the_header_types():
    header = _header_host_forms()
    decls = _ffi_decls()
    assert len(header) > 80
    for name, params in header.items():
        assert name in decls, f"{name} missing from ffi.rs"
        got, ret = decls[name]
        assert ret == "i32", name
        assert got == params, f"{name}: ffi.rs has {got}, header has {params}"
    for extra in ("coral_init", "coral_last_error", "coral_status_string"):
        assert extra in decls
    assert set(n for n in decls if n.startswith("coralh_")) == set(header)


def test_lib_binds_coral_surface_module_by_module():
    text = _read("src", "lib.rs")
    decls = _ffi_decls()
    for mod, names in (("level1", LEVEL1), ("level2", LEVEL2), ("level3", LEVEL3), ("split", SPLIT)):
        bound = _bound(_module(text, mod))
        for name in names:
            assert bound.get(name) == f"coralh_{name}", f"{mod}::{name} is not bound to coralh_{name}"
            assert f"coralh_{name}" in decls
    assert len(LEVEL1) + len(LEVEL2) + len(LEVEL3) == 80


def test_enums_keep_the_c_abi_codes():
    text = _module(_read("src", "lib.rs"), "enums")
    order = {e: re.findall(r"(\w+),", body)
             for e, body in re.findall(r"pub enum (\w+) \{(.*?)\}", text, flags=re.S)}
    assert order["CoralTranspose"] == ["NoTranspose", "Transpose", "ConjugateTranspose"]
    assert order["CoralTriangular"] == ["UpperTriangular", "LowerTriangular"]
    assert order["CoralDiagonal"] == ["UnitDiagonal", "NonUnitDiagonal"]
    with open(_lib.HEADER_PATH) as f:
        h = f.read()
    for name, code in (("CORAL_NO_TRANSPOSE", 0), ("CORAL_TRANSPOSE", 1), ("CORAL_CONJUGATE_TRANSPOSE", 2),
                       ("CORAL_UPPER_TRIANGULAR", 0), ("CORAL_LOWER_TRIANGULAR", 1),
                       ("CORAL_UNIT_DIAGONAL", 0), ("CORAL_NON_UNIT_DIAGONAL", 1)):
        assert re.search(rf"\b{name}\s*=\s*{code}\b", h), name


def test_crate_links_the_library():
    assert 'name = "coral_hip"' in _read("Cargo.toml")
    build = _read("build.rs")
    assert "rustc-link-lib=dylib=coralhip" in build and "CORAL_HIP_LIB_DIR" in build
