"""Locates (and, when missing or stale, builds with make) the C++ case
programs of tests/cpp, which link the in-tree libfocalors_poisson.so."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CPP = os.path.join(HERE, "cpp")
OUT = os.path.join(CPP, "_build")


def program(name):
    r = subprocess.run(["make", "-s", "-C", CPP, os.path.join(OUT, name)], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("building %s failed:\n%s" % (name, r.stderr[-4000:]))
    return os.path.join(OUT, name)


def run(name, *args, timeout=120):
    r = subprocess.run([program(name), *map(str, args)], capture_output=True, text=True, timeout=timeout)
    if r.returncode != 0:
        raise AssertionError("%s exited %d:\n%s\n%s" % (name, r.returncode, r.stdout[-2000:], r.stderr[-4000:]))
    return r.stdout
