"""Every `<reference file>:<line>[-<line>]` citation in this repository's
sources and documents names a file of the reference and lines that exist in
it (line counts in tests/golden/reference_lines.json).  A citation of a
bare file name must resolve to exactly one reference file unless all
candidates are long enough."""
import json
import re
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
TABLE = json.loads((ROOT / "tests/golden/reference_lines.json").read_text())
SKIP_DIRS = {"_work", ".git", "gpurun_out", "node_modules", "__pycache__", "_build", "build", "_ref"}
# Documents written for this repository by others, not sources of it.
SKIP_FILES = {"VERDICT.md", "ADVICE.md", "SURVEY.md", "BASELINE.md"}
SUFFIXES = {".py", ".c", ".cpp", ".hpp", ".h", ".hip", ".js", ".ts", ".md", ".txt"}
CITE = re.compile(r"((?:[\w.-]+/)*[\w.-]+\.(?:ts|glsl|frag|vert|js|json|yml|md)):(\d+)(?:\s*[-–]\s*(\d+))?"
                  r"((?:,\s*\d+(?:\s*[-–]\s*\d+)?)*)")


def candidates(name: str):
    name = name.removeprefix("/root/reference/").lstrip("./")
    if name in TABLE:
        return [name]
    return [p for p in TABLE if p == name or p.endswith("/" + name)]


def citations():
    for p in sorted(ROOT.rglob("*")):
        rel = p.relative_to(ROOT)
        if not p.is_file() or p.suffix not in SUFFIXES or SKIP_DIRS & set(rel.parts) or p.name in SKIP_FILES:
            continue
        for ln, line in enumerate(p.read_text(errors="replace").splitlines(), 1):
            for m in CITE.finditer(line):
                ranges = [(int(m.group(2)), int(m.group(3) or m.group(2)))]
                for extra in re.finditer(r"(\d+)(?:\s*[-–]\s*(\d+))?", m.group(4) or ""):
                    ranges.append((int(extra.group(1)), int(extra.group(2) or extra.group(1))))
                yield f"{rel}:{ln}", m.group(1), ranges


def test_table_covers_the_reference_sources():
    assert len(TABLE) > 40 and "src/renderer.ts" in TABLE and "src/shaders/trace.frag" in TABLE


def test_every_citation_points_at_existing_lines():
    bad, checked = [], 0
    for where, name, ranges in citations():
        cands = candidates(name)
        if not cands:
            # A file of this repository (bench.py's JSON, node/lucent.js, DESIGN.md ...) is not a reference citation.
            if list(ROOT.rglob(Path(name).name)):
                continue
            bad.append(f"{where}: {name} is not a reference file")
            continue
        checked += 1
        longest = max(TABLE[c] for c in cands)
        for a, b in ranges:
            if not (1 <= a <= b <= longest):
                bad.append(f"{where}: {name}:{a}-{b} outside 1..{longest}")
    assert checked > 300
    assert not bad, "\n".join(bad)


def test_the_checker_catches_a_bad_citation(tmp_path):
    """The pattern and the bounds check reject a line past a file's end."""
    line = "see src/shaders/bsdf.glsl" + ":186-406 and trace.frag" + ":53-272, 324-999"
    found = [(m.group(1), m.group(2), m.group(3), m.group(4)) for m in CITE.finditer(line)]
    assert found == [("src/shaders/bsdf.glsl", "186", "406", ""), ("trace.frag", "53", "272", ", 324-999")]
    assert TABLE[candidates("bsdf.glsl")[0]] < 406
    assert TABLE[candidates("trace.frag")[0]] < 999
