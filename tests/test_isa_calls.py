"""Code-generation guard for the HIP library (CPU only: compiles to gfx950
assembly and reads it).

A device function that is called (not inlined) keeps its return address in
s[30:31].  When such a function is longer than the short-branch range, the
compiler's long-branch expansion (s_getpc_b64 / s_add_u32 / s_setpc_b64 to a
local label) may use that same register pair, and the function's final
s_setpc_b64 s[30:31] then jumps back into its own body instead of returning:
the kernel never ends.  The triangular-dislocation code is therefore written
to be inlined and short; this test checks every translation unit's device
assembly for long-branch expansions inside called functions.
"""
import os
import re
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "quake-dfn_amd", "csrc")
HIPCC = "/opt/rocm/bin/hipcc"
FLAGS = ["-std=c++17", "-O3", "--offload-arch=gfx950", "-ffp-contract=off", "-Xclang", "-target-feature",
         "-Xclang", "-packed-fp32-ops", "--cuda-device-only", "-S", f"-I{os.path.join(ROOT, 'include')}",
         f"-I{CSRC}"]


def _asm(src, out_dir):
    out = os.path.join(out_dir, os.path.basename(src) + ".s")
    subprocess.run([HIPCC, *FLAGS, src, "-o", out], check=True, capture_output=True, timeout=600)
    return open(out).read()


def _functions(asm):
    """(name, is_kernel, body lines) for every function in the listing."""
    kernels = set(re.findall(r"^\s*\.amdhsa_kernel\s+(\S+)", asm, re.M))
    funcs = []
    cur, body = None, []
    for line in asm.split("\n"):
        m = re.match(r"^([A-Za-z_][\w.$]*):", line)
        if m and not m.group(1).startswith("."):
            if cur:
                funcs.append((cur, cur in kernels, body))
            cur, body = m.group(1), []
            continue
        if cur is not None:
            if line.startswith(".Lfunc_end"):
                funcs.append((cur, cur in kernels, body))
                cur, body = None, []
            else:
                body.append(line)
    return funcs


def _long_branches(body):
    """Indices of s_getpc_b64 whose address is advanced to a local label."""
    hits = []
    for i, line in enumerate(body):
        if "s_getpc_b64" in line:
            window = " ".join(body[i + 1:i + 4])
            if re.search(r"\(\.LBB\d+_\d+-", window):
                hits.append(i)
    return hits


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_no_long_branch_in_called_functions(tmp_path):
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        listings = list(ex.map(lambda s: (s, _asm(s, str(tmp_path))), srcs))
    bad = []
    n_kernels = 0
    for src, asm in listings:
        for name, is_kernel, body in _functions(asm):
            n_kernels += is_kernel
            if not is_kernel and _long_branches(body):
                bad.append(f"{os.path.basename(src)}: {name}")
    assert n_kernels >= 10
    assert not bad, "long-branch expansion inside called device functions: " + ", ".join(bad)
