"""The index arithmetic of a captured batch step replayed over inputs whose
per-mesh vote counts differ from the captured ones, checked on the CPU with
AddressSanitizer and UndefinedBehaviorSanitizer (tests/replay_index_check.cpp:
segments after the running-maximum clamp, segment lookup, cell keys and their
width, cell offsets, query row ranges, stratified seed draws, all with the
captured step's buffer sizes).

The current lookup keeps every index inside its buffer over hundreds of
random shrunk, grown and unchanged count sets.  The lookup the fault was
made with (points past the last segment read the grid of the segment after
the last) is reported by AddressSanitizer as a read past the grid array:
that read is how a replay over fewer votes corrupted the cell keys, the
sort and the offsets that the query's row ranges come from."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "replay_index_check.cpp")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if not shutil.which("g++"):
        pytest.skip("no host C++ compiler")
    exe = str(tmp_path_factory.mktemp("replay") / "replay_index_check")
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined",
                        "-fno-sanitize-recover=all", "-o", exe, SRC],
                       capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip("sanitizer build unavailable: " + r.stderr[-300:])
    return exe


@pytest.mark.parametrize("seed", [1, 2])
def test_changed_counts_keep_indices_in_bounds(checker, seed):
    r = subprocess.run([checker, str(seed), "250"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    assert r.stdout.startswith("ok 250")


def test_old_segment_lookup_reads_past_the_grids(checker):
    r = subprocess.run([checker, "1", "250", "old"], capture_output=True, text=True, timeout=300)
    assert r.returncode != 0
    assert "heap-buffer-overflow" in r.stderr or "key is wider" in r.stderr, r.stderr[-2000:]
