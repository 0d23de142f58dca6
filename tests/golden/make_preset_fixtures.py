"""Writes tests/golden/preset_frames_fp64.json: SHA-256 digests of the frame
files that the reference's own sequential driver (benchapps/nbody/seq/nbody_main.cc,
FTYPE double, compiled unmodified into oracle/_ref/nbody_main_seq by
oracle/Makefile) writes for every simulation preset (-i 0..6) at several N,
two steps each: `nbody_main_seq -n N -i P -l 2 -O out.txt`.

The preset table and the N-dependent cluster / velocity scales
(seq/nbody_main.cc:57-66,109-144) set timestep, softening and the initial
conditions, so each digest pins one row of the table at one N branch.  Every
N is a multiple of 256, so the UTC driver (which rounds N up to a multiple of
its block size) runs the same system.  N = 33024 lies above the last
threshold, where the preset's own scales apply.

Run in the build container (the reference tree must be present and
oracle/_ref built):  python tests/golden/make_preset_fixtures.py
"""
from __future__ import annotations

import concurrent.futures as cf
import hashlib
import json
import os
import subprocess
import tempfile
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
SEQ = ROOT / "oracle" / "_ref" / "nbody_main_seq"
OUT = HERE / "preset_frames_fp64.json"
PRESETS = range(7)
SIZES = [1024, 2048, 4096, 8192, 16384, 32768, 33024]
STEPS = 2


def one(preset: int, n: int) -> dict:
    with tempfile.TemporaryDirectory() as d:
        out = Path(d) / "out.txt"
        cmd = [str(SEQ), "-n", str(n), "-i", str(preset), "-l", str(STEPS), "-O", str(out)]
        subprocess.run(cmd, cwd=d, check=True, capture_output=True, timeout=1200)
        data = out.read_bytes()
    lines = data.decode().splitlines()
    return {"preset": preset, "n": n, "steps": STEPS, "bytes": len(data),
            "sha256": hashlib.sha256(data).hexdigest(),
            "frame1_label": lines[n + 1], "frame1_body0": lines[n + 2]}


def main() -> None:
    if not SEQ.exists():
        raise SystemExit(f"{SEQ} missing: build oracle/_ref first (make -C oracle)")
    jobs = [(p, n) for n in sorted(SIZES, reverse=True) for p in PRESETS]
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        res = list(ex.map(lambda a: one(*a), jobs))
    res.sort(key=lambda r: (r["n"], r["preset"]))
    OUT.write_text(json.dumps({
        "source": "benchapps/nbody/seq/nbody_main.cc (FTYPE double), built by oracle/Makefile",
        "command": "nbody_main_seq -n N -i PRESET -l 2 -O out.txt",
        "frames": res}, indent=1) + "\n")
    print(f"wrote {OUT} ({len(res)} runs)")


if __name__ == "__main__":
    main()
