"""The C-ABI library loads and exports every symbol include/nwick.h declares."""
import os
import re

import nwick

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_header_symbols_exported():
    hdr = open(os.path.join(ROOT, "include", "nwick.h")).read()
    declared = set(re.findall(r"\b(nw_[a-z0-9_]+)\s*\(", hdr))
    assert declared == set(nwick.EXPORTS)
    L = nwick.load()
    for name in declared:
        assert hasattr(L, name)


def test_det_record_layout():
    assert nwick.DET_DTYPE.itemsize == 120
    assert nwick.DET_DTYPE.fields["oa"][1] == 24
    assert nwick.DET_DTYPE.fields["pb"][1] == 104
