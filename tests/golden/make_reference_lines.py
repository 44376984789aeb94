"""Writes reference_lines.json: the line count of every text file of the
reference repository (TypeScript, GLSL, JavaScript, JSON, YAML, Markdown),
keyed by its path relative to the repository root.  tests/test_citations.py
checks every `<file>:<line>` citation in this repository against it.

    python tests/golden/make_reference_lines.py [/root/reference]
"""
import json
import sys
from pathlib import Path

TEXT = {".ts", ".js", ".glsl", ".frag", ".vert", ".json", ".yml", ".md", ".html", ".css"}


def main():
    ref = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    table = {}
    for p in sorted(ref.rglob("*")):
        rel = p.relative_to(ref)
        if p.is_file() and p.suffix in TEXT and not {".git", "node_modules"} & set(rel.parts):
            table[rel.as_posix()] = len(p.read_text(encoding="utf-8", errors="replace").splitlines())
    out = Path(__file__).resolve().parent / "reference_lines.json"
    out.write_text(json.dumps(table, indent=1, sort_keys=True) + "\n")
    print(f"{len(table)} files -> {out}")


if __name__ == "__main__":
    main()
