"""Copy the reference's own nbody golden frame files into tests/golden/.

Both files are data the reference ships beside its sources (2 frames x 1024
bodies, "%f" frame label then "%.5f %.5f %.5f %.5f" per body):
  benchapps/nbody/seq/nbody_outpos.txt   sequential CPU path, FP32 build,
                                         `-n 1024 -l 100` (preset 0, SHELL)
  benchapps/nbody/cuda/nbody_outpos.txt  the reference's CUDA path, same run
Run once in a container that has the reference tree; the copies are committed
so the GPU tests can read them where the reference is absent.
"""
import hashlib
import shutil
from pathlib import Path

SRC = Path("/root/reference/utc-0.1/benchapps/nbody")
DST = Path(__file__).resolve().parent
FILES = {
    "seq/nbody_outpos.txt": ("seq_nbody_outpos.txt",
                             "1484534f18e875397249e74b1a319161a7b0d70307ee4dffc1ea1683082ce1d3"),
    "cuda/nbody_outpos.txt": ("cuda_nbody_outpos.txt",
                              "005a22f2839ceb88237345a1b3c954853295d3a11d72a3e1a7376395f5223ccc"),
}

if __name__ == "__main__":
    for rel, (name, sha) in FILES.items():
        data = (SRC / rel).read_bytes()
        assert hashlib.sha256(data).hexdigest() == sha, rel
        shutil.copyfile(SRC / rel, DST / name)
        print("copied", rel, "->", name)
