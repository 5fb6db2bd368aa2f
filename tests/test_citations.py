"""Every reference citation in the product, the oracle, the tests and the
design notes resolves in the reference tree.

A citation is ``File.ext:L`` or ``File.ext:L1-L2[,L3-L4...]`` naming one of
the reference's source files.  The test checks that the file exists, that
every cited line lies inside it, and, where a citation directly follows a
routine name (``update_cv_and_bv (TimeAdvanceFunc.cu:255-290)``), that the
routine's name occurs inside the cited range.  It reads the reference
tree only as text and is skipped where that tree is absent.
"""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.environ.get("CFX_REFERENCE_TREE", "/root/reference")

CITE = re.compile(r"(\bsrc/)?\b([A-Z][A-Za-z]*\.(?:cu|cuh|cpp|h))(?::(\d+(?:-\d+)?(?:,\s?\d+(?:-\d+)?)*))?")
# a routine name immediately before "(File:lines" or "File:lines" after "name ("
NAMED = re.compile(r"([A-Za-z_][A-Za-z_0-9]*(?:::[A-Za-z_][A-Za-z_0-9]*)?)`? \(`?([A-Z][A-Za-z]*\.(?:cu|cuh|cpp|h)):(\d+)")
SCANNED = ("include", "oracle", "combustioncodegpu_amd", "tests", "DESIGN.md", "INTEGRATION.md", "bench.py",
           "__graft_entry__.py", "README.md")
SELF = os.path.abspath(__file__)


def _sources():
    out = []
    for r in SCANNED:
        p = os.path.join(ROOT, r)
        if os.path.isfile(p):
            out.append(p)
            continue
        for dp, dn, fn in os.walk(p):
            dn[:] = [d for d in dn if d not in ("_build", "__pycache__", "golden", "_ref")]
            for f in fn:
                if f.endswith((".h", ".cpp", ".hip", ".py", ".md")) and os.path.join(dp, f) != SELF:
                    out.append(os.path.join(dp, f))
    return sorted(out)


@pytest.fixture(scope="module")
def reference():
    if not os.path.isdir(REF):
        pytest.skip("reference tree not present")
    files = {}
    for f in os.listdir(REF):
        p = os.path.join(REF, f)
        if os.path.isfile(p):
            with open(p, errors="replace") as fh:
                files[f] = fh.read().split("\n")
    return files


def test_citations_resolve(reference):
    bad, n = [], 0
    for p in _sources():
        with open(p, errors="replace") as fh:
            for ln, line in enumerate(fh, 1):
                for m in CITE.finditer(line):
                    prefix, name, rng = m.groups()
                    if name not in reference:
                        continue
                    n += 1
                    where = f"{os.path.relpath(p, ROOT)}:{ln}: {m.group(0)}"
                    if prefix:
                        bad.append(where + " (the reference has no src/ directory)")
                    if rng:
                        last = len(reference[name]) - (reference[name][-1] == "")
                        nums = [int(x) for x in re.findall(r"\d+", rng)]
                        if min(nums) < 1 or max(nums) > last:
                            bad.append(f"{where} (file has {last} lines)")
    assert n > 150, f"only {n} citations found; the scanner lost track of the sources"
    assert not bad, "citations outside their file:\n" + "\n".join(bad)


def test_named_citations_cover_the_routine(reference):
    """A routine name written right before a citation (snake_case or
    Class::member) appears inside the cited line range."""
    bad, n = [], 0
    for p in _sources():
        with open(p, errors="replace") as fh:
            for ln, line in enumerate(fh, 1):
                for m in NAMED.finditer(line):
                    routine, name = m.group(1), m.group(2)
                    if name not in reference or not (re.fullmatch(r"[a-z][a-z0-9]*_[a-z0-9_]+", routine) or "::" in routine):
                        continue
                    rng = line[m.end(3):].split(")")[0]
                    first = int(m.group(3))
                    last = first
                    tail = re.match(r"-(\d+)", rng)
                    if tail:
                        last = int(tail.group(1))
                    if last > len(reference[name]):
                        continue
                    n += 1
                    short = routine.split("::")[-1]
                    body = "\n".join(reference[name][first - 1:last])
                    if short not in body:
                        bad.append(f"{os.path.relpath(p, ROOT)}:{ln}: {routine} -> {name}:{first}-{last}: "
                                   f"{reference[name][first - 1].strip()[:70]}")
    assert n > 20, f"only {n} named citations checked"
    assert not bad, "named citations whose range does not hold the routine:\n" + "\n".join(bad)
