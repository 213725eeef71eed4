"""The ctypes mirror of ``struct vd_qg_params`` and ``struct vd_obs_table`` in
vardyn_amd/_lib.py lists the header's fields in the header's order (a
mismatch would make the library read the wrong words)."""
import re
from pathlib import Path

from vardyn_amd import _lib

HEADER = Path(__file__).resolve().parents[1] / "include" / "vardyn.h"


def struct_fields(name):
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    body = re.search(r"typedef struct %s \{(.*?)\}" % name, text, flags=re.S).group(1)
    fields = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        names = decl.split(None, 1)[1] if not decl.startswith("const") else decl.split(None, 2)[2]
        for n in names.split(","):
            fields.append(n.strip().lstrip("*").strip())
    return fields


def test_params_layout_matches_header():
    assert [f for f, _ in _lib.QgParams._fields_] == struct_fields("vd_qg_params")


def test_obs_table_layout_matches_header():
    assert [f for f, _ in _lib.ObsTable._fields_] == struct_fields("vd_obs_table")


def test_bm_params_layout_matches_header():
    from vardyn_amd.basis_bm import BmParams
    assert [f for f, _ in BmParams._fields_] == struct_fields("vd_bm_params")


def test_merge_params_layout_matches_header():
    from vardyn_amd.merge import MergeParams
    assert [f for f, _ in MergeParams._fields_] == struct_fields("vd_merge_params")


def test_sw_params_layout_matches_header():
    assert [f for f, _ in _lib.SwParams._fields_] == struct_fields("vd_sw_params")
