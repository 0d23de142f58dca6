"""CPU checks of bench.py's contract pieces that need no GPU: it refuses more
than one GPU with a message, and the HBM traffic it reports is read from the
committed PMC summary."""
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_traffic_comes_from_the_committed_pmc_summary():
    import bench
    t = bench.read_traffic(ROOT / "profiles" / "force_mixed_n262144_pmc.txt")
    assert t is not None and 1.5e9 < t < 2.5e9
    assert bench.read_traffic(ROOT / "profiles" / "no-such-file.txt") is None
