"""The host code of the library under AddressSanitizer + UBSan and under
ThreadSanitizer (CPU only): argument errors, footprints and the host layout,
including multi-particle jobs laid out on several threads, raise no report."""
import pytest

from peridem_amd import sanitize


@pytest.mark.parametrize("kind", ["asan", "tsan"])
def test_host_code_is_clean_under_the_sanitizer(kind):
    if sanitize.runtime(kind) is None:
        pytest.skip(f"no {kind} runtime in this toolchain")
    tests, layout = sanitize.run(kind)
    assert tests.returncode == 0, tests.stdout[-3000:] + tests.stderr[-3000:]
    assert layout.returncode == 0 and "layout ok" in layout.stdout, layout.stderr[-3000:]
