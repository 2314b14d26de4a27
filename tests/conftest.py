import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def _rss_guard(limit_gib):
    """Ends the test process, naming the running test, when its resident set
    passes limit_gib: a runaway host allocation then fails one test run
    instead of exhausting the machine."""
    import threading
    import time

    def run():
        while True:
            time.sleep(0.5)
            try:
                with open("/proc/self/status") as fh:
                    kb = next(int(l.split()[1]) for l in fh if l.startswith("VmRSS:"))
            except (OSError, StopIteration, ValueError):
                return
            if kb > limit_gib * 1024 * 1024:
                msg = "\nresident memory %.1f GiB above the %g GiB guard in %s\n" % (
                    kb / 2.0 ** 20, limit_gib, os.environ.get("PYTEST_CURRENT_TEST", "?"))
                os.write(sys.__stderr__.fileno() if sys.__stderr__ else 2, msg.encode())
                os._exit(3)

    threading.Thread(target=run, name="rss-guard", daemon=True).start()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")
    if sys.platform.startswith("linux"):
        _rss_guard(float(os.environ.get("FP_TEST_RSS_LIMIT_GIB", "96")))
