"""Command line and mesh files without a GPU: the demo generator against the
reference's example files, and the CLI's argument, file and mesh contracts
that fail before the device is touched (the reference's tests/pipeline.py
error checks and src/app/main.cpp, src/io/mesh.cpp messages)."""
import filecmp
import os
import subprocess

import pytest

from mitra_algo_amd import demo

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "mitra_algo_amd", "mvs_cli")
REF_CLI = os.path.join(ROOT, "oracle", "_ref", "ref_cli")
GOLDEN = os.path.join(ROOT, "tests", "golden")
needs_cli = pytest.mark.skipif(not os.path.exists(CLI), reason="mvs_cli not built")


def run(*args, ok=False):
    p = subprocess.run([CLI, *map(str, args)], capture_output=True, text=True, timeout=60)
    assert (p.returncode == 0) == ok, p.stdout + p.stderr
    return p


@pytest.mark.parametrize("kind", ["reflection", "rigid", "partial", "single"])
def test_demo_reproduces_reference_examples(tmp_path, kind):
    out = tmp_path / f"{kind}.off"
    demo.write_off(out, *demo.demo(kind))
    assert filecmp.cmp(out, os.path.join(GOLDEN, f"{kind}.off"), shallow=False)


def test_demo_shape_counts():
    v, f = demo.shape(0)
    assert len(v) == 6 and len(f) == 8
    for level in (1, 2, 3):
        v, f = demo.shape(level)
        # Euler: a closed genus-0 surface with F = 8 * 4^level faces
        assert len(f) == 8 * 4 ** level and len(v) - 3 * len(f) // 2 + len(f) == 2


@needs_cli
def test_usage_and_help():
    p = subprocess.run([CLI], capture_output=True, text=True, timeout=60)
    assert p.returncode == 2 and "Usage" in p.stdout
    assert run("--help", ok=True).stdout.startswith("Partial symmetry detector")


@needs_cli
def test_argument_errors(tmp_path):
    model = tmp_path / "m.off"
    demo.write_off(model, *demo.demo("reflection"))
    assert "Input and --output are required" in run(model).stderr
    assert "Only one input mesh" in run(model, model, "--output", tmp_path / "o").stderr
    assert "Missing value for --seed" in run(model, "--output", tmp_path / "o", "--seed").stderr
    assert "Unknown option: --bogus" in run(model, "--output", tmp_path / "o", "--bogus",
                                           "1").stderr
    assert "Unknown backend" in run(model, "--output", tmp_path / "o", "--backend",
                                    "bogus").stderr
    assert "CPU backend unavailable" in run(model, "--output", tmp_path / "o", "--backend",
                                            "cpu").stderr
    for option, value, msg in [("--sample-ratio", "0", "Invalid parameter range"),
                               ("--bandwidth", "nan", "Invalid numeric value"),
                               ("--smoothing", "-1", "Invalid integer"),
                               ("--prune", "1", "Invalid parameter range"),
                               ("--seed", "-1", "Invalid unsigned"),
                               ("--max-clusters", "1.5", "Invalid integer"),
                               ("--max-pairs", "4294967296", "32-bit ID capacity")]:
        assert msg in run(model, "--output", tmp_path / "invalid", option, value).stderr
    # existing results are preserved
    (tmp_path / "full").mkdir()
    (tmp_path / "full" / "x").write_text("keep")
    assert "new or empty" in run(model, "--output", tmp_path / "full").stderr
    assert (tmp_path / "full" / "x").read_text() == "keep"


@needs_cli
def test_mesh_file_errors(tmp_path):
    bad = tmp_path / "bad.off"
    bad.write_text("OFF\n3 1 0\n0 0 0\n1 0 0\n0 1 0\n3 0 1 9\n")
    assert "range" in run(bad, "--output", tmp_path / "a").stderr
    bad.write_text("OFF\n3 1 0\n0 0 0\n1 0 0\n2 0 0\n3 0 1 2\n")
    assert "nondegenerate" in run(bad, "--output", tmp_path / "b").stderr
    v, f = demo.demo("reflection")
    f = list(f)
    a, b, c = f[0]
    f[0] = (a, c, b)
    demo.write_off(bad, v, f)
    assert "winding" in run(bad, "--output", tmp_path / "c").stderr
    bad.write_text("OFF 3\n")
    assert "Invalid OFF header" in run(bad, "--output", tmp_path / "d").stderr
    bad.write_text("OFF\n3 1 0\n0 0 0\n1 0 0\n")
    assert "Truncated OFF vertices" in run(bad, "--output", tmp_path / "e").stderr
    bad.write_text("OFF\n3 1 0\n0 0 0\n1 0 0\n0 1 0\n2 0 1\n")
    assert "Invalid OFF polygon" in run(bad, "--output", tmp_path / "f").stderr
    # the stream parse refuses "nan", as the reference's does
    bad.write_text("OFF\n3 1 0\n0 0 0\n1 0 0\n0 1 nan\n3 0 1 2\n")
    assert "Truncated OFF vertices" in run(bad, "--output", tmp_path / "g").stderr
    obj = tmp_path / "m.obj"
    obj.write_text("v 0 0 0\nv 1 0 0\nv 1 1 0\nf 1 2 5\n")
    assert "OBJ index out of range" in run(obj, "--output", tmp_path / "h").stderr
    obj.write_text("v 0 0 0\nv 1 0 0\nv 1 1 0\nf 1 0 2\n")
    assert "Invalid OBJ index" in run(obj, "--output", tmp_path / "i").stderr
    obj.write_text("v 0 0 0\nv 1 0 0\nf 1 2\n")
    assert "Invalid OBJ face" in run(obj, "--output", tmp_path / "j").stderr
    ply = tmp_path / "m.ply"
    ply.write_text("ply\n")
    assert "Supported mesh formats" in run(ply, "--output", tmp_path / "k").stderr
    assert "Cannot open mesh" in run(tmp_path / "none.off", "--output", tmp_path / "l").stderr


BAD_FILES = {
    "range.off": "OFF\n3 1 0\n0 0 0\n1 0 0\n0 1 0\n3 0 1 9\n",
    "degenerate.off": "OFF\n3 1 0\n0 0 0\n1 0 0\n2 0 0\n3 0 1 2\n",
    "header.off": "OFF 3\n",
    "short.off": "OFF\n3 1 0\n0 0 0\n1 0 0\n",
    "polygon.off": "OFF\n3 1 0\n0 0 0\n1 0 0\n0 1 0\n2 0 1\n",
    "nan.off": "OFF\n3 1 0\n0 0 0\n1 0 0\n0 1 nan\n3 0 1 2\n",
    "comment.off": "OFF # c\n3 1 0 # counts\n0 0 0\n1 0 0\n0 0 0\n3 0 1 2 # face\n",
    "range.obj": "v 0 0 0\nv 1 0 0\nv 1 1 0\nf 1 2 5\n",
    "zero.obj": "v 0 0 0\nv 1 0 0\nv 1 1 0\nf 1 0 2\n",
    "text.obj": "v 0 0 0\nv 1 0 0\nv 1 1 0\nf 1 2 x\n",
    "short.obj": "v 0 0 0\nv 1 0 0\nf 1 2\n",
    "vertex.obj": "v 0 0\nf 1 2 3\n",
    "empty.obj": "# nothing\n",
    "m.ply": "ply\n",
}


@needs_cli
@pytest.mark.skipif(not os.path.exists(REF_CLI), reason="oracle/_ref/ref_cli not built")
@pytest.mark.parametrize("name", sorted(BAD_FILES))
def test_mesh_errors_match_reference_cli(tmp_path, name):
    path = tmp_path / name
    path.write_text(BAD_FILES[name])
    ours = run(path, "--output", tmp_path / "ours")
    ref = subprocess.run([REF_CLI, str(path), "--output", str(tmp_path / "ref")],
                         capture_output=True, text=True, timeout=60)
    assert ref.returncode == 1
    assert ours.stderr == ref.stderr


@needs_cli
@pytest.mark.skipif(not os.path.exists(REF_CLI), reason="oracle/_ref/ref_cli not built")
@pytest.mark.parametrize("args", [["--sample-ratio", "0"], ["--bandwidth", "nan"],
                                  ["--smoothing", "-1"], ["--prune", "1"], ["--seed", "-1"],
                                  ["--max-clusters", "1.5"], ["--max-pairs", "4294967296"],
                                  ["--verify-threshold", "0.5"], ["--subdivide", "4"],
                                  ["--max-iterations", "0"], ["--bogus", "1"], ["--seed"]])
def test_option_errors_match_reference_cli(tmp_path, args):
    model = tmp_path / "m.off"
    demo.write_off(model, *demo.demo("single"))
    ours = run(model, "--output", tmp_path / "ours", *args)
    ref = subprocess.run([REF_CLI, str(model), "--output", str(tmp_path / "ref"), *args],
                         capture_output=True, text=True, timeout=60)
    assert ref.returncode == 1
    assert ours.stderr == ref.stderr
