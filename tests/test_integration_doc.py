"""The C binding listed in INTEGRATION.md compiles against the shipped
headers (so the listing and the ABI cannot drift apart)."""
import os
import re
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_integration_listing_compiles(tmp_path):
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = re.findall(r"```c\n(.*?)```", text, re.S)
    assert len(blocks) == 1
    src = tmp_path / "semlb_pfg.c"
    src.write_text(blocks[0])
    r = subprocess.run(["gcc", "-std=c11", "-Wall", "-Wextra", "-Werror", "-Wno-unused-function",
                        "-I", os.path.join(ROOT, "include"), "-c", str(src), "-o",
                        str(tmp_path / "semlb_pfg.o")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
