"""The host driver's threads -- worker pool, results saver, progress reporter
(the reference's ThreadedTaskProcessor / WorkerThread / ResultsSaver /
ProgressReporter) -- with a stub solver on the CPU, built and run under
ThreadSanitizer: tasks are computed and saved exactly once whatever the number
of workers, a worker's exception reaches the caller, and TSAN reports no data
race (it exits with status 66 when it does)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "pincosimulator_amd", "host")


def test_worker_pool_saver_and_reporter_under_tsan(tmp_path):
    exe = str(tmp_path / "task_processor_check")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=thread", "-I", HOST,
           os.path.join(ROOT, "tests", "native", "task_processor_check.cpp"),
           os.path.join(HOST, "task_processor.cpp"), os.path.join(HOST, "pinco_format.cpp"),
           "-o", exe, "-lpthread"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1 exitcode=66")
    r = subprocess.run([exe, str(tmp_path / "out")], capture_output=True, text=True, env=env, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr[-4000:]
    assert "ALL OK" in r.stdout
    assert "WARNING: ThreadSanitizer" not in r.stderr, r.stderr[-4000:]
