"""Every `path:line` / `path:first-last` citation of the reference in the C-ABI header, the package's headers
and sources, the oracle and the tests names a file of the reference tree and lines that exist in it.

Runs only where the reference tree is present (G1_REFERENCE_ROOT, default /root/reference); it is read as
text, nothing of it is built or imported. A citation may be relative to the reference root, to its
`include/` or `include/GraphOne/` directories, or to the directory of the citation before it in the same
file (as in `archi/blas/spmv/def.hpp:8-73, cpu/naive.hpp:5-20`); a bare file name must be unique in the tree.
"""
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.environ.get("G1_REFERENCE_ROOT", "/root/reference")
CITE = re.compile(r"([A-Za-z0-9_./-]+\.(?:h|hpp|cuh|cu|cpp|txt|md)):(\d+)(?:-(\d+))?")

pytestmark = pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "include", "GraphOne")),
                                reason="reference tree not present")


def _sources():
    pats = ["include/*.h", "graphone_amd/include/**/*.h", "graphone_amd/include/**/*.hpp",
            "graphone_amd/csrc/*.hip", "graphone_amd/csrc/*.cpp", "graphone_amd/csrc/*.hpp",
            "graphone_amd/*.py", "graphone_amd/apps/*.cpp", "oracle/*.c", "oracle/*.cpp", "oracle/*.py",
            "tests/*.py", "tests/golden/*.py", "bench.py", "__graft_entry__.py"]
    out = []
    for p in pats:
        out += glob.glob(os.path.join(ROOT, p), recursive=True)
    return sorted(set(out))


def _ref_index():
    idx = {}
    for dp, _, fns in os.walk(REF):
        for f in fns:
            idx.setdefault(f, []).append(os.path.relpath(os.path.join(dp, f), REF))
    return idx


def _resolve(path, lastdir, idx):
    for base in ("", "include/", "include/GraphOne/"):
        if os.path.isfile(os.path.join(REF, base + path)):
            return base + path
    if lastdir and os.path.isfile(os.path.join(REF, lastdir, path)):
        return os.path.join(lastdir, path)
    hits = [p for p in idx.get(os.path.basename(path), []) if p.endswith(path)]
    return hits[0] if len(hits) == 1 else None


def test_every_reference_citation_is_in_range():
    idx = _ref_index()
    lines = {}
    bad = []
    checked = 0
    for src in _sources():
        if os.path.basename(src) == "test_citations.py":
            continue
        text = open(src, errors="ignore").read()
        lastdir = None
        for m in CITE.finditer(text):
            path, first = m.group(1), int(m.group(2))
            last = int(m.group(3) or first)
            rel = _resolve(path, lastdir, idx)
            where = f"{os.path.relpath(src, ROOT)}: {m.group(0)}"
            if rel is None:
                bad.append(f"{where} (no such file in the reference)")
                continue
            lastdir = os.path.dirname(rel)
            if rel not in lines:
                with open(os.path.join(REF, rel), errors="ignore") as f:
                    lines[rel] = sum(1 for _ in f)
            checked += 1
            if not 1 <= first <= last <= lines[rel]:
                bad.append(f"{where} ({rel} has {lines[rel]} lines)")
    assert checked > 100
    assert not bad, "\n".join(bad)
