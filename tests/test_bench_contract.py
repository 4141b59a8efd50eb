"""The benchmark's command-line contract that holds without a GPU: a run on
more than one GPU is refused with a message and exit status 2 (the path is
built for one GPU), also when launched as one rank of a multi-rank job, and
before anything touches the device."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run_bench(args, env_extra=None):
    env = dict(os.environ)
    env.update(env_extra or {})
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py")] + args, cwd=ROOT, env=env,
                          capture_output=True, text=True, timeout=300)


def test_more_than_one_gpu_is_refused():
    r = run_bench(["--gpus", "2", "--steps", "1", "--warmup", "0"])
    assert r.returncode == 2, (r.returncode, r.stdout, r.stderr)
    assert "one GPU only" in r.stderr
    assert r.stdout.strip() == ""  # no JSON line for a refused run


def test_rank_of_a_multi_rank_job_is_refused():
    r = run_bench(["--steps", "1", "--warmup", "0"], {"WORLD_SIZE": "4", "RANK": "1", "LOCAL_RANK": "1"})
    assert r.returncode == 2, (r.returncode, r.stdout, r.stderr)
    assert "one GPU only" in r.stderr
