"""Builds copies of the oracle with one planted slip each (text replacements
in oracle/kpax_oracle.cpp), for the tests that show a slip is caught."""
import os
import shutil
import subprocess
import tempfile
from concurrent.futures import ThreadPoolExecutor

from conftest import ROOT

SRC = os.path.join(ROOT, "oracle", "kpax_oracle.cpp")


def sources(replacements):
    """[(what, source)] with each (what, text, replacement) applied alone; the
    text must occur exactly once."""
    src = open(SRC).read()
    out = []
    for what, a, b in replacements:
        assert src.count(a) == 1, f"slip '{what}': its text occurs {src.count(a)} times in the oracle"
        out.append((what, src.replace(a, b)))
    return out


class Variants:
    """Context manager: builds every source into its own shared library in a
    temporary directory; yields [(what, path)]."""

    def __init__(self, srcs):
        self.srcs = srcs

    def __enter__(self):
        if shutil.which("g++") is None:
            import pytest
            pytest.skip("no g++")
        self.tmp = tempfile.mkdtemp(prefix="kpax_oracle_variants_")
        inc = os.path.join(ROOT, "include")

        def build(i):
            cpp = os.path.join(self.tmp, f"v{i}.cpp")
            with open(cpp, "w") as f:
                f.write(self.srcs[i][1].replace('#include "../include/kpax.h"', '#include "kpax.h"'))
            so = os.path.join(self.tmp, f"libv{i}.so")
            r = subprocess.run(["g++", "-std=c++17", "-O1", "-ffp-contract=off", "-mfma", "-fopenmp", "-fPIC",
                                "-shared", "-I", inc, cpp, "-o", so], capture_output=True, text=True)
            assert r.returncode == 0, r.stderr[-2000:]
            return so

        with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
            libs = list(ex.map(build, range(len(self.srcs))))
        return [(what, so) for (what, _), so in zip(self.srcs, libs)]

    def __exit__(self, *a):
        shutil.rmtree(self.tmp, ignore_errors=True)
