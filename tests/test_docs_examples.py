"""The C example of INTEGRATION.md compiles against include/plbpd_hip.h and
links against the built library (no GPU call: compile and link only)."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "plbpd_amd", "lib", "libplbpd_hip.so")


def c_example():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    section = text[text.index("## 3. The C ABI directly"):]
    return re.search(r"```c\n(.*?)```", section, re.S).group(1)


@pytest.mark.skipif(shutil.which("gcc") is None or not os.path.exists(LIB), reason="needs gcc and the built library")
def test_integration_c_example_compiles_and_links(tmp_path):
    body = "\n".join(l for l in c_example().splitlines() if not l.startswith("#include"))
    src = tmp_path / "example.c"
    src.write_text("#include <stdint.h>\n#include <stdio.h>\n#include <stdlib.h>\n#include \"plbpd_hip.h\"\n\n"
                   "int example(const uint8_t* solid_mask, double* u_host) {\n" + body + "    return 0;\n}\n")
    out = tmp_path / "libexample.so"
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-shared", "-fPIC", str(src), "-I",
                        os.path.join(ROOT, "include"), "-L", os.path.dirname(LIB), "-lplbpd_hip",
                        "-Wl,--no-undefined", "-o", str(out)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def python_example():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    section = text[text.index("## 4. Python (ctypes) binding"):]
    return re.search(r"```python\n(.*?)```", section, re.S).group(1)


@pytest.mark.skipif(not os.path.exists(LIB), reason="needs the built library")
def test_integration_ctypes_stub_runs(monkeypatch):
    """The ctypes stub runs as written: it creates and destroys a state on a
    GPU, and without one fails through its own check() with the library's
    no-device message."""
    monkeypatch.chdir(ROOT)
    try:
        exec(compile(python_example(), "INTEGRATION.md", "exec"), {})
    except RuntimeError as e:
        assert "no HIP device" in str(e)
