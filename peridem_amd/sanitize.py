"""Sanitizer builds of the library's host code, and a run of its host-only
paths under them.

The HIP device code is built as usual; only the host part carries the
sanitizer (each -fsanitize= goes to the host pass through -Xarch_host):

  asan   AddressSanitizer + UndefinedBehaviorSanitizer
  tsan   ThreadSanitizer

Each build lands in peridem_amd/_san/<kind>/libperidem_gpu.so.  `run` loads it
in a fresh Python (PERIDEM_GPU_LIB, the sanitizer runtime preloaded) and runs
what needs no GPU: tests/test_prepare.py (argument errors, footprints, the
host layout) and the layout of multi-particle aggregate jobs, whose particles
pd_job_prepare lays out on up to 16 threads.

    python -m peridem_amd.sanitize asan      # or tsan
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "csrc", "pd_api.hip")

KINDS = {
    "asan": (["-fsanitize=address", "-fsanitize=undefined", "-fno-omit-frame-pointer"],
             "libclang_rt.asan-x86_64.so",
             {"ASAN_OPTIONS": "detect_leaks=0:abort_on_error=1",
              "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1"}),
    "tsan": (["-fsanitize=thread"], "libclang_rt.tsan-x86_64.so",
             {"TSAN_OPTIONS": "halt_on_error=1:report_signal_unsafe=0"}),
}

LAYOUT = """
import sys
sys.path.insert(0, {root!r})
from peridem_amd import abi, gen
assert abi.GPU_LIB == {lib!r}, abi.GPU_LIB
for k in (5, 1, 7):
    b = abi.prepare_job(gen.aggregate_job(k))
    b.footprint()
    b.free()
print("layout ok")
"""


def runtime(kind):
    r = subprocess.run(["hipcc", f"-print-file-name={KINDS[kind][1]}"], capture_output=True, text=True)
    path = r.stdout.strip()
    return path if r.returncode == 0 and os.path.isabs(path) and os.path.exists(path) else None


def build(kind):
    flags, _, _ = KINDS[kind]
    out_dir = os.path.join(HERE, "_san", kind)
    os.makedirs(out_dir, exist_ok=True)
    out = os.path.join(out_dir, "libperidem_gpu.so")
    cmd = ["hipcc", "-O1", "-g", "-std=c++17", "-fPIC", "-shared", "-pthread", "-ffp-contract=off",
           "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-Wno-unused-result", "-Wno-unused-value"]
    for f in flags:
        cmd += ["-Xarch_host", f]
    subprocess.run(cmd + ["-o", out, SRC], check=True, capture_output=True, text=True)
    return out


def run(kind):
    """Build, then run the host-only paths under the sanitizer; returns the
    completed processes (a sanitizer report makes them fail)."""
    lib = build(kind)
    env = dict(os.environ)
    env.update(KINDS[kind][2])
    env["PERIDEM_GPU_LIB"] = lib
    env["LD_PRELOAD"] = runtime(kind)
    tests = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider",
                            os.path.join(ROOT, "tests", "test_prepare.py")],
                           env=env, capture_output=True, text=True, timeout=900)
    layout = subprocess.run([sys.executable, "-c", LAYOUT.format(root=ROOT, lib=lib)], env=env,
                            capture_output=True, text=True, timeout=900)
    return tests, layout


if __name__ == "__main__":
    kind = sys.argv[1] if len(sys.argv) > 1 else "asan"
    t, l = run(kind)
    sys.stdout.write(t.stdout[-2000:] + t.stderr[-4000:] + l.stdout + l.stderr[-4000:])
    sys.exit(t.returncode or l.returncode)
