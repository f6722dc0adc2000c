"""Every HIP kernel of the built library runs without scratch (private)
memory: a spill or a register array indexed by a lane-dependent value moves
data to scratch, a slow path the kernels are written to avoid.  Reads the
kernel metadata (.private_segment_fixed_size) of the library's own gfx950
code objects with the ROCm LLVM tools; no GPU needed."""
import os
import re
import shutil
import subprocess
import tempfile

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(REPO, "ucfd_sparse_amd", "liblusgs.so")
LLVM = "/opt/rocm/llvm/bin"
MAGIC = b"__CLANG_OFFLOAD_BUNDLE__"


def tools():
    b = os.path.join(LLVM, "clang-offload-bundler")
    r = os.path.join(LLVM, "llvm-readelf")
    if not (os.path.exists(b) and os.path.exists(r) and shutil.which("objcopy")):
        pytest.skip("ROCm LLVM tools absent")
    return b, r


def kernel_metadata():
    """(kernel name, private segment bytes) of every kernel in the library."""
    bundler, readelf = tools()
    if not os.path.exists(LIB):
        pytest.skip("library not built")
    out = []
    with tempfile.TemporaryDirectory() as d:
        fat = os.path.join(d, "fatbin")
        # an explicit output file: without one objcopy rewrites its input
        subprocess.run(["objcopy", "--dump-section", f".hip_fatbin={fat}", LIB,
                        os.path.join(d, "copy.so")], check=True)
        data = open(fat, "rb").read()
        starts = [m.start() for m in re.finditer(re.escape(MAGIC), data)]
        for k, s in enumerate(starts):
            blob = os.path.join(d, f"b{k}")
            open(blob, "wb").write(data[s:starts[k + 1] if k + 1 < len(starts) else len(data)])
            co = os.path.join(d, f"co{k}.o")
            subprocess.run([bundler, "--unbundle", "--type=o", f"--input={blob}",
                            "--targets=hipv4-amdgcn-amd-amdhsa--gfx950", f"--output={co}"],
                           check=True)
            notes = subprocess.run([readelf, "--notes", co], check=True, capture_output=True,
                                   text=True).stdout
            name = None
            for line in notes.splitlines():
                line = line.strip()
                if line.startswith(".name:"):
                    name = line.split(":", 1)[1].strip()
                elif line.startswith(".private_segment_fixed_size:"):
                    out.append((name, int(line.split(":", 1)[1])))
    return out


def test_every_kernel_runs_without_scratch():
    meta = kernel_metadata()
    assert len(meta) > 100, len(meta)
    spills = [(n, b) for n, b in meta if b]
    assert not spills, spills
