"""The reference's callers of the FDK path (cuFDK/mpiFDK.cpp, cudaBP/cudaBP.cpp,
FilterlibCufft/FilterlibCufft.cpp, genProj/genProj.cpp) compile against the C++
drop-in header, in their own call forms; tests/callsites_check.cpp holds those
forms.  Host compiler only, no GPU."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
@pytest.mark.parametrize("std", ["c++11", "c++17"])
def test_reference_call_sites_compile(std):
    r = subprocess.run(["g++", f"-std={std}", "-Wall", "-Wextra", "-Werror", "-fsyntax-only",
                        "-I", os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "callsites_check.cpp")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
