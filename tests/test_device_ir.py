"""Guard against a miscompile of the device code by this image's hipcc.

hipcc (ROCm 7.2) once compiled the probe kernel's filter target
`op == FILTER_KNOWN ? row[ec] : desc.kend` into a pointer phi whose
known_to_const input structurizecfg replaced by poison; on the GPU those
lanes loaded from a stale register and faulted the device.  This test
compiles the kernels with the library's device flags, prints the IR right
after control-flow structurization, and fails if any global load or store
takes its address from a pointer phi that can be poison.  It runs on the CPU
(hipcc cross-compiles for gfx950) and needs no GPU.
"""
import os
import re
import subprocess

import pytest

from wukong_gpu import _build

CLANG = "/opt/rocm/lib/llvm/bin/clang++"


def _pointer_phis(func_text):
    phis = {}
    for ln in func_text.split("\n"):
        m = re.match(r"\s+(%[\w.]+) = phi ptr(?: addrspace\(\d\))? (.*)", ln)
        if m:
            phis[m.group(1)] = re.findall(r"\[ (%[\w.]+|poison|undef|null), %[\w.]+ \]", m.group(2))
    return phis


def poisoned_memory_ops(ir_text):
    """(function, instruction) for every load/store whose pointer operand is a
    phi chain with a poison or undef incoming value."""
    bad = []
    for f in re.split(r"(?m)^define ", ir_text)[1:]:
        name = re.match(r".*?@(\S+)\(", f).group(1)
        phis = _pointer_phis(f)

        def reaches(v, seen=()):
            if v in ("poison", "undef"):
                return True
            if v not in phis or v in seen:
                return False
            return any(reaches(x, seen + (v,)) for x in phis[v])

        for ln in f.split("\n"):
            m = (re.search(r"= load [^,]+, ptr(?: addrspace\(\d\))? (%[\w.]+)", ln)
                 or re.search(r"store [^,]+, ptr(?: addrspace\(\d\))? (%[\w.]+)", ln))
            if m and reaches(m.group(1)):
                bad.append((name, ln.strip()))
    return bad


def test_checker_flags_the_known_bad_shape():
    """The scan recognises the shape the miscompile produced."""
    ir = """define void @f(ptr addrspace(1) %a, i1 %c) {
1:
  br i1 %c, label %2, label %3
2:
  %p = getelementptr i8, ptr addrspace(1) %a, i64 24
  br label %3
3:
  %q = phi ptr addrspace(1) [ %p, %2 ], [ poison, %1 ]
  %v = load i32, ptr addrspace(1) %q, align 4
  ret void
}
"""
    assert len(poisoned_memory_ops(ir)) == 1
    assert poisoned_memory_ops(ir.replace("poison", "%a")) == []


@pytest.mark.skipif(not os.path.exists(CLANG), reason="ROCm clang not present")
def test_structurized_kernels_have_no_poisoned_addresses(tmp_path):
    src = os.path.join(_build.CSRC, "kernels.hip")
    cmd = ([CLANG, "-x", "hip", "--cuda-device-only", "-c", src, "-o", os.devnull]
           + [f for f in _build.COMMON if f != "-fopenmp"] + _build.DEVICE
           + ["-mllvm", "-print-after=amdgpu-annotate-uniform"])
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    ir = r.stderr
    kernels = re.findall(r"(?m)^define amdgpu_kernel void @(\S+)\(", ir)
    assert any("k_probe" in k for k in kernels) and any("k_gather" in k for k in kernels), kernels
    assert poisoned_memory_ops(ir) == []
