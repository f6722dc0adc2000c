"""The C-ABI boundary without a GPU: liblusgs.so loads, exports every function
the public headers declare, and a plain C caller compiles and links against
the headers exactly as the reference examples call the legacy API.  No
compute entry point is called here."""
import os
import re
import subprocess
import tempfile

import pytest

from ucfd_sparse_amd import lib as L

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(REPO, "include")


def declared_functions():
    names = {}
    for fn in sorted(os.listdir(INC)):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(INC, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        text = re.sub(r"#[^\n]*", "", text)
        text = re.sub(r"typedef[^;]*;", "", text)
        for m in re.finditer(r"\b([A-Za-z_]\w*)\s*\(([^;{}]*)\)\s*;", text):
            name = m.group(1)
            if name in ("if", "while", "for", "return", "sizeof"):
                continue
            names.setdefault(fn, []).append(name)
    return names


def test_headers_declare_the_legacy_api():
    names = {n for v in declared_functions().values() for n in v}
    for n in ("parallel_pre_lusgs", "parallel_ns_lower_sweep", "parallel_ns_upper_sweep",
              "ns_parallel_lower_sweep", "ns_parallel_upper_sweep", "lusgs_parallel_ns_update",
              "serial_pre_lusgs", "serial_ns_lower_sweep", "serial_ns_upper_sweep",
              "ns_serial_lower_sweep", "ns_serial_upper_sweep", "serial_ns_lusgs_update",
              "lusgs_serial_ns_update"):
        assert n in names


def test_library_exports_every_declared_symbol():
    lib = L.load()
    missing = [(h, n) for h, ns in declared_functions().items() for n in ns
               if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_every_declared_symbol_has_ctypes_argtypes():
    """The Python binding never calls an entry point with default ctypes
    conversions (which would truncate 64-bit pointers)."""
    lib = L.load()
    unbound = [n for ns in declared_functions().values() for n in ns
               if getattr(lib, n).argtypes is None]
    assert not unbound, unbound


def test_legacy_signatures_match_reference_ctypes_bindings():
    """argtypes of examples/omp3d.py:27-35 and examples/mpi3d.py:26-34."""
    import ctypes
    lib = L.load()
    vp, ci, cd = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    assert lib.ns_parallel_lower_sweep.argtypes == [ci] * 4 + [vp] * 10
    assert lib.parallel_pre_lusgs.argtypes == [ci, ci, cd] + [vp] * 4
    assert lib.ns_serial_upper_sweep.argtypes == [ci, ci] + [vp] * 8
    assert lib.lusgs_serial_ns_update.argtypes == [ci, vp, vp]


@pytest.mark.parametrize("prog", ["legacy_calls", "krylov_calls"])
def test_c_caller_compiles_and_links(prog):
    src = os.path.join(REPO, "tests", "abi", prog + ".c")
    libdir = os.path.join(REPO, "ucfd_sparse_amd")
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "a.x")
        r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", INC, src, "-o", exe,
                            "-L", libdir, "-llusgs", "-lm", f"-Wl,-rpath,{libdir}"],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr


def test_headers_declare_the_krylov_api():
    """Every public entry point of src/krylov/include/ucfd{mat,precon,solver}.h
    that the CPU/CUDA builds offer, under the same names."""
    names = {n for v in declared_functions().values() for n in v}
    for n in ("UCFDMatInit", "UCFDMatDestroy", "UCFDMatMult", "UCFDMatCreateCSR",
              "UCFDMatCreateBSR", "UCFDMatCreateCUDACSR", "UCFDMatCreateCUDABSR",
              "UCFDPreconCreatefromArrays", "UCFDPreconCreateNone", "UCFDPreconPrepare",
              "UCFDPreconDestroy", "UCFDPreconSetILU", "UCFDPreconSetBILU",
              "UCFDPreconSetBLUSGS", "UCFDPreconSetPBILU", "UCFDPreconSetCUDABLUSGS",
              "UCFDPreconSetCUDABILU",
              "UCFDSolverInit", "UCFDSolverDestroy", "UCFDSolve", "UCFDSolverGetResult",
              "UCFDSolverTraceResidualHistory", "UCFDSolverGetResidualHistory",
              "UCFDSolverCreateGMRES", "UCFDSolverCreateBICGSTAB", "UCFDSolverCreateCUDAGMRES"):
        assert n in names, n


def test_headers_declare_the_flowsys_api():
    """Every entry point of src/lusgs/include/ucfdlusgs.h:8-40."""
    names = {n for v in declared_functions().values() for n in v}
    for n in ("UCFDFlowSysCreate", "UCFDFlowSysSetElement", "UCFDFlowSysDestroy",
              "UCFDFlowSysSetLUSGS", "UCFDLUSGS_Pack", "UCFDLUSGS_Update", "UCFDLUSGS_NSPrepare",
              "UCFDLUSGS_RANSPrepare", "UCFDLUSGS_NSLowerSweep", "UCFDLUSGS_NSUpperSweep",
              "UCFDLUSGS_RANSLowerSweep", "UCFDLUSGS_RANSUpperSweep", "UCFDFlowSysSetBLUSGS",
              "UCFDBLUSGS_Pack", "UCFDBLUSGS_KWSST_Pack", "UCFDBLUSGS_SA_Pack",
              "UCFDBLUSGS_Update", "UCFDBLUSGS_SubResidual", "UCFDBLUSGS_NSPrepare",
              "UCFDBLUSGS_RANSPrepare", "UCFDBLUSGS_NSLowerSweep", "UCFDBLUSGS_RANSLowerSweep",
              "UCFDBLUSGS_NSUpperSweep", "UCFDBLUSGS_RANSUpperSweep", "UCFDBLUSGS_Reset"):
        assert n in names, n


def test_empty_legacy_calls_return_before_touching_arrays():
    """Zero cells, or an empty colour range [n0, n0), is a successful no-op in
    the reference's loops; here too, before any array is read (NULL arrays
    allowed), so the calls need no GPU."""
    for name, args in (
            ("parallel_pre_lusgs", (0, 4, 1.0, None, None, None, None)),
            ("serial_pre_lusgs", (0, 6, 1.0, None, None, None, None)),
            ("parallel_ns_lower_sweep", (3, 3, 10, 4) + (None,) * 9),
            ("parallel_ns_upper_sweep", (0, 0, 0, 4) + (None,) * 9),
            ("ns_parallel_lower_sweep", (7, 7, 10, 4) + (None,) * 10),
            ("ns_parallel_upper_sweep", (7, 7, 10, 4) + (None,) * 10),
            ("serial_ns_lower_sweep", (0, 4) + (None,) * 7),
            ("serial_ns_upper_sweep", (0, 4) + (None,) * 7),
            ("ns_serial_lower_sweep", (0, 4) + (None,) * 8),
            ("ns_serial_upper_sweep", (0, 4) + (None,) * 8),
            ("lusgs_parallel_ns_update", (0, None, None)),
            ("lusgs_serial_ns_update", (0, None, None)),
            ("serial_ns_lusgs_update", (0, None, None))):
        assert L.call(name, *args) == 0, name


def test_device_count_call_is_safe_without_gpu():
    assert L.device_count() >= 0
