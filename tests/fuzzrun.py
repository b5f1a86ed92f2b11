"""Runs one golden fuzz case through a tn93-compatible binary (the gate of evaluate.py:61-99)."""
import os
import subprocess
from pathlib import Path


def run_case(binary, case, td: Path, env=None, timeout=60):
    files = {k: v.encode("latin-1") for k, v in case["files"].items()}
    names = {k: f"{k}.fa" for k in files}
    for k, data in files.items():
        (td / names[k]).write_bytes(data)
    argv = [str(binary), "-q"] + [names[f.strip("{}")] if f.startswith("{") else f for f in case["flags"]]
    argv += ["-o", "out.txt", names["in"]]
    out = td / "out.txt"
    if out.exists():
        out.unlink()
    e = dict(os.environ)
    if env:
        e.update(env)
    p = subprocess.run(argv, cwd=td, capture_output=True, timeout=timeout, env=e)
    return p.returncode, (out.read_bytes() if out.exists() else None), p.stdout, p.stderr


def check_case(binary, case, td, env=None):
    """Returns None when the binary matches the fixture, else a short description of the difference."""
    rc, out, so, err = run_case(binary, case, td, env)
    if case["rc"] != 0:
        return None if rc != 0 else f"expected failure (rc {case['rc']}), got rc 0"
    if rc != 0:
        return f"rc {rc}: {err[-300:]!r}"
    exp_out = None if case["out"] is None else case["out"].encode("latin-1")
    if out != exp_out:
        return f"output file differs: {(out or b'')[:200]!r} vs {(exp_out or b'')[:200]!r}"
    if so != case["stdout"].encode("latin-1"):
        return f"stdout differs: {so[:300]!r} vs {case['stdout'][:300]!r}"
    return None
