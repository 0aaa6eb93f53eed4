"""bench.py's command-line contract on the CPU: it measures one GPU, so more than one GPU (by
flag or by a launcher's WORLD_SIZE) exits with status 2 and says so, before touching a GPU."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run_bench(args, env_extra=None):
    env = dict(os.environ)
    env.update(env_extra or {})
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py")] + args, cwd=ROOT, env=env,
                          capture_output=True, text=True, timeout=120)


def test_more_than_one_gpu_exits():
    for args, extra in [(["--gpus", "2"], None), (["--gpus", "8", "--steps", "1"], None),
                        ([], {"WORLD_SIZE": "2"})]:
        r = run_bench(args, extra)
        assert r.returncode == 2, (args, extra, r.stderr)
        assert "one GPU" in r.stderr
        assert r.stdout == ""  # no JSON line


def test_help_lists_the_legs_and_switches():
    r = run_bench(["--help"])
    assert r.returncode == 0
    for flag in ("--gpus", "--steps", "--warmup", "--no-cpu", "--no-large", "--no-serve", "--no-harness",
                 "--calibrate-copy", "--serial-cosets"):
        assert flag in r.stdout, flag
