"""The Julia binding in INTEGRATION.md against the C ABI: every struct the
stub passes through ccall has the fields of its include/fsi_api.h
counterpart, with the same names, types and order, and every entry point it
calls is declared there (no Julia toolchain exists here, so this is the
check that the stub would bind the library as declared)."""
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
HEADER = (ROOT / "include" / "fsi_api.h").read_text()
DOC = (ROOT / "INTEGRATION.md").read_text()

JULIA_TO_C = {
    "BcSide": "fsi_bc_side", "BcSetup": "fsi_bc_setup", "Bcs": "fsi_bcs", "Grid": "fsi_grid",
    "Models": "fsi_models", "Numerics": "fsi_numerics", "CStepReport": "fsi_step_report",
    "StepStats": "fsi_step_stats", "ProjectionResultC": "fsi_projection_result", "DiagParams": "fsi_diag_params",
    "TensorRegion": "fsi_tensor_region", "Diagnostics": "fsi_diagnostics", "AdjointStats": "fsi_adjoint_stats",
}
SCALARS = {"Int32": "int32_t", "Int64": "int64_t", "Float64": "double"}


def _julia_block():
    m = re.search(r"```julia\n(.*?)```", DOC, re.S)
    assert m, "no julia block in INTEGRATION.md"
    return m.group(1)


def _julia_structs(code):
    out = {}
    for m in re.finditer(r"^struct (\w+)(.*?)\bend\b", code, re.S | re.M):
        name, body = m.group(1), m.group(2)
        body = re.sub(r"#[^\n]*", "", body)
        out[name] = re.findall(r"(\w+)::([\w{},]+)", body)
    return out


def _c_structs():
    text = re.sub(r"/\*.*?\*/", "", HEADER, flags=re.S)
    macros = {k: int(v) for k, v in re.findall(r"#define (\w+) (\d+)", HEADER)}
    out = {}
    for m in re.finditer(r"typedef struct \{(.*?)\} (\w+);", text, re.S):
        fields = []
        for decl in m.group(1).split(";"):
            decl = decl.strip()
            if not decl:
                continue
            ty, names = decl.split(None, 1)
            for n in names.split(","):
                n = n.strip()
                am = re.match(r"(\w+)\[(\w+)\]", n)
                if am:
                    size = am.group(2)
                    fields.append((am.group(1), ty, int(macros.get(size, size))))
                else:
                    fields.append((n, ty, None))
        out[m.group(2)] = fields
    return out


def _julia_type(t):
    """(C scalar or struct type, array length or None) of a Julia field type."""
    m = re.match(r"NTuple\{(\d+),(\w+)\}", t)
    if m:
        return _julia_type(m.group(2))[0], int(m.group(1))
    if t in SCALARS:
        return SCALARS[t], None
    if t in JULIA_TO_C:
        return JULIA_TO_C[t], None
    raise AssertionError(f"unmapped Julia type {t}")


@pytest.mark.parametrize("jname", sorted(JULIA_TO_C))
def test_stub_struct_matches_header(jname):
    js = _julia_structs(_julia_block())
    cs = _c_structs()
    assert jname in js, f"struct {jname} missing from the Julia stub"
    cname = JULIA_TO_C[jname]
    assert cname in cs, f"{cname} missing from fsi_api.h"
    jf = [(n, *_julia_type(t)) for n, t in js[jname]]
    cf = cs[cname]
    # C's int32_t fields hold the enums and booleans; the names follow the header
    assert [f[0] for f in jf] == [f[0] for f in cf], (jname, [f[0] for f in jf], [f[0] for f in cf])
    assert jf == cf, (jname, jf, cf)


def test_stub_calls_only_declared_entry_points():
    code = _julia_block()
    called = set(re.findall(r"ccall\(\(:(\w+)", code)) | set(re.findall(r":(fsi_\w+)\)", code))
    declared = set(re.findall(r"\b(fsi_\w+)\(", re.sub(r"/\*.*?\*/", "", HEADER, flags=re.S)))
    assert called, "no ccall found"
    missing = sorted(c for c in called if c not in declared)
    assert not missing, missing
    for needed in ("fsi_context_create", "fsi_workspace_create", "fsi_state_create", "fsi_integrate_step",
                   "fsi_make_velocity_compatible", "fsi_state_download", "fsi_state_upload"):
        assert needed in called, needed


def test_stub_keeps_the_reference_step_signature():
    """integrate_step!(state, t, dt, ctx, work) with five positional
    arguments (src/TimeIntegrationIMEX.jl:1152-1158)."""
    code = _julia_block()
    m = re.search(r"function EulerianFSI\.integrate_step!\(([^)]*)\)", code)
    assert m
    args = [a.strip() for a in m.group(1).split(",")]
    assert len(args) == 5 and args[0].startswith("state") and args[3].startswith("ctx") and \
        args[4].startswith("work"), args
    for helper in ("function models_struct", "function numerics_struct", "function unpack_diagnostics"):
        assert helper in code, helper
