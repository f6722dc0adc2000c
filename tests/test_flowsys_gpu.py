"""FlowSys block and scalar LU-SGS on the GPU, through the C ABI (include/ucfdlusgs.h),
against the CPU oracle on the same seeded inputs: every stage of an outer
step (pack, turbulence pack, NS / RANS prepare, reset, lower / upper sweeps,
sub-residual, update) bit for bit, on a randomly numbered mesh (many
dependency levels) and on a colour-ordered one (one launch per colour), with
host and device arrays.  The FP32 sweep form stays within 1e-5 of FP64."""
import os
import sys

import numpy as np
import pytest

from oracle import pyoracle as O
from ucfd_sparse_amd import mesh as M

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def say(msg):
    print(f"[flowsys-gpu] {msg}", flush=True)


def case(m, seed, colored=False, nvars=6):
    from ucfd_sparse_amd.layout import colored_layout
    mesh = M.tet_mesh(m, seed=seed, permute=True)
    flow = M.flow_fields(mesh, seed=seed)
    if colored:
        lay = colored_layout(mesh, flow)
        mesh, flow = lay.mesh, lay.flow
    fs = M.flowsys_system(mesh, flow, nv=5, seed=seed)
    tj = M.flowsys_system(mesh, flow, nv=nvars - 5, seed=seed).jmat
    el = M.flowsys_elements(mesh, flow, nvars=nvars, seed=seed)
    return fs, tj, el


def on_device(fs, tj, el):
    from ucfd_sparse_amd.mesh import FlowSysMesh

    def d(a):
        return torch.from_numpy(np.ascontiguousarray(a)).cuda()
    fsd = FlowSysMesh(fs.nlocal, fs.nfaces, d(fs.rowptr), d(fs.colidx), d(fs.slots),
                      d(fs.sides), d(fs.face_area), d(fs.face_normal), d(fs.rcp_vol), d(fs.jmat))
    eld = {k: d(v) for k, v in el.items()}
    return fsd, d(tj), eld


@pytest.mark.parametrize("where,colored,kind", [("host", False, "sa"), ("device", False, "sa"),
                                                ("device", True, "sa"),
                                                ("device", True, "kwsst")])
def test_stages_bitwise(where, colored, kind):
    from ucfd_sparse_amd import flowsys as FS
    nvars = 7 if kind == "kwsst" else 6
    say(f"arrays on {where}, colour-ordered={colored}, turbulence pack {kind}")
    fs, tj, el = case(4, 71, colored, nvars)
    ref = O.flowsys_run(fs, tj, el, nvars, 5, kind=kind, iterations=2)
    if where == "device":
        fsd, tjd, eld = on_device(fs, tj, el)
        S = FS.FlowSys(fsd, nvars, 5, tjmat=tjd)
        S.set_element(0, eld)
    else:
        S = FS.FlowSys(fs, nvars, 5, tjmat=tj)
        S.set_element(0, el)
    got = FS.run(S, kind=kind, iterations=2)
    S.close()
    for k, v in ref.items():
        g = np.asarray(got[k]).reshape(v.shape)
        assert np.array_equal(g, v), k


def split_slots(fs, tj, seed, nsplit, nsame, nmerge=0, nempty=0):
    """The same system with nsplit faces split into two one-position slots
    (each with a copy of the face's Jacobians, area and normal: boundary-like
    slots), nsame faces whose two positions carry the same side, nmerge
    faces whose second position moves into another face's slot (a slot of
    three positions and one of one), and nempty extra slots that no position
    names."""
    from ucfd_sparse_amd.mesh import FlowSysMesh
    rng = np.random.default_rng(seed)
    slots, sides = fs.slots.copy(), fs.sides.copy()
    order = np.argsort(slots, kind="stable")
    pairs = order.reshape(-1, 2)            # every generated slot holds two positions
    pick = rng.choice(len(pairs), size=nsplit + nsame + 2*nmerge, replace=False)
    ns = fs.nfaces
    new_f = []
    for j, k in enumerate(pick[:nsplit]):
        p, q = pairs[k]
        new_f.append(slots[q])
        slots[q] = ns + j
    for k in pick[nsplit:nsplit + nsame]:
        p, q = pairs[k]
        sides[q] = sides[p]
    m = pick[nsplit + nsame:]
    for k1, k2 in zip(m[0::2], m[1::2]):
        slots[pairs[k2][1]] = slots[pairs[k1][0]]
    new_f += [int(f) for f in rng.choice(ns, size=nempty)]
    nsplit += nempty
    new_f = np.array(new_f, np.int64)
    jm = np.concatenate([fs.jmat, fs.jmat[..., new_f]], axis=-1)
    tjm = np.concatenate([tj, tj[..., new_f]], axis=-1)
    fa = np.concatenate([fs.face_area, fs.face_area[new_f]])
    fn = np.concatenate([fs.face_normal, fs.face_normal[:, new_f]], axis=1)
    return FlowSysMesh(fs.nlocal, ns + nsplit, fs.rowptr, fs.colidx, slots, sides, fa, fn,
                       fs.rcp_vol, np.ascontiguousarray(jm)), np.ascontiguousarray(tjm)


@pytest.mark.parametrize("colored", [False, True])
def test_stages_bitwise_with_unpaired_slots(colored):
    """Prepare takes a position's diagonal contribution from its partner's
    packed block when the slot holds exactly two positions of opposite sides,
    and from jmat otherwise: slots with one position, a pair with equal
    sides and slots of three positions exercise the second path; slots that
    no position names are skipped by the slot-ordered packing."""
    from ucfd_sparse_amd import flowsys as FS
    say(f"unpaired and same-side slots, colour-ordered={colored}")
    fs, tj, el = case(4, 73, colored, 6)
    fs, tj = split_slots(fs, tj, 74, nsplit=40, nsame=3, nmerge=5, nempty=7)
    ref = O.flowsys_run(fs, tj, el, 6, 5, kind="sa", iterations=2)
    fsd, tjd, eld = on_device(fs, tj, el)
    S = FS.FlowSys(fsd, 6, 5, tjmat=tjd)
    S.set_element(0, eld)
    got = FS.run(S, kind="sa", iterations=2)
    S.close()
    for k, v in ref.items():
        g = np.asarray(got[k]).reshape(v.shape)
        assert np.array_equal(g, v), k


@pytest.mark.parametrize("nvars,nfvars", [(7, 7), (9, 7)])
def test_hex_wide_blocks_bitwise(nvars, nfvars):
    """7x7 blocks on a colour-ordered hex mesh: one wave's positions exceed the
    LDS staging of the wave kernels (their HBM-direct path), and a 2x2
    turbulence block for nvars = 9."""
    from ucfd_sparse_amd import flowsys as FS
    from ucfd_sparse_amd.layout import colored_layout
    say(f"hex mesh, nvars={nvars}, nfvars={nfvars}")
    mesh = M.hex_mesh(7, 6, 5, seed=81, permute=True)
    flow = M.flow_fields(mesh, seed=81)
    lay = colored_layout(mesh, flow, coloring="greedy-bfs")
    mesh, flow = lay.mesh, lay.flow
    fs = M.flowsys_system(mesh, flow, nv=nfvars, seed=81)
    tj = M.flowsys_system(mesh, flow, nv=nvars - nfvars, seed=81).jmat if nvars > nfvars else None
    el = M.flowsys_elements(mesh, flow, nvars=nvars, seed=81)
    ref = O.flowsys_run(fs, tj, el, nvars, nfvars, kind="kwsst", iterations=2)
    fsd, tjd, eld = on_device(fs, tj if tj is not None else np.zeros(1), el)
    S = FS.FlowSys(fsd, nvars, nfvars, tjmat=tjd if tj is not None else None)
    S.set_element(0, eld)
    got = FS.run(S, kind="kwsst", iterations=2)
    S.close()
    for k, v in ref.items():
        g = np.asarray(got[k]).reshape(v.shape)
        assert np.array_equal(g, v), k


def test_fp32_sweeps_close_to_fp64():
    from ucfd_sparse_amd import flowsys as FS
    say("FP32 sweep products vs FP64")
    fs, tj, el = case(6, 72, colored=True)
    fsd, tjd, eld = on_device(fs, tj, el)
    out = {}
    for fp32 in (False, True):
        S = FS.FlowSys(fsd, 6, 5, tjmat=tjd, fp32=fp32)
        S.set_element(0, {k: v.clone() for k, v in eld.items()})
        out[fp32] = FS.run(S, iterations=3)
        S.close()
    for k in ("du_upper0", "du_upper2", "du_rans2"):
        a, b = out[False][k], out[True][k]
        err = np.linalg.norm(a - b) / np.linalg.norm(a)
        say(f"  {k}: rel err {err:.2e}")
        assert 0 < err <= 1e-5


def test_state_and_index_errors():
    from ucfd_sparse_amd import flowsys as FS
    from ucfd_sparse_amd import lib as L
    say("error paths")
    fs, tj, el = case(3, 73)
    S = FS.FlowSys(fs, 6, 5, tjmat=tj)
    S.set_element(0, el)
    with pytest.raises(L.UCFDError):
        S.call("UCFDBLUSGS_NSLowerSweep")            # before Prepare
    S.close()
    bad = fs.colidx.copy()
    bad[3] = fs.nlocal + 2
    from ucfd_sparse_amd.mesh import FlowSysMesh
    fsb = FlowSysMesh(fs.nlocal, fs.nfaces, fs.rowptr, bad, fs.slots, fs.sides, fs.face_area,
                      fs.face_normal, fs.rcp_vol, fs.jmat)
    S = FS.FlowSys(fsb, 6, 5, tjmat=tj)
    S.set_element(0, el)
    S.call("UCFDBLUSGS_Pack", 0, 0.3)
    with pytest.raises(L.UCFDError):
        S.call("UCFDBLUSGS_NSPrepare")
    S.close()
    elb = dict(el, cell_ids=el["cell_ids"].copy())
    elb["cell_ids"][5] = -1
    S = FS.FlowSys(fs, 6, 5, tjmat=tj)
    S.set_element(0, elb)
    with pytest.raises(L.UCFDError):
        S.call("UCFDBLUSGS_Pack", 0, 0.3)
    S.close()
    # the library keeps working
    S = FS.FlowSys(fs, 6, 5, tjmat=tj)
    S.set_element(0, el)
    got = FS.run(S, iterations=1)
    ref = O.flowsys_run(fs, tj, el, 6, 5, iterations=1)
    assert np.array_equal(got["du_upper0"], ref["du_upper0"])
    S.close()


# ---------------------------------------------------------------- scalar LU-SGS
def lusgs_case(seed=304, m=3, colored=False):
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
    import make_golden
    return make_golden.flowsys_lusgs_case(seed=seed, m=m, colored=colored)


@pytest.mark.parametrize("where,colored,m", [("host", False, 3), ("device", False, 3),
                                             ("device", True, 4)])
def test_lusgs_stages_bitwise(where, colored, m):
    from ucfd_sparse_amd import flowsys as FS
    say(f"scalar LU-SGS, arrays on {where}, colour-ordered={colored}, m={m}")
    mesh, flow, fs, fspr, tfspr, el = lusgs_case(seed=304 + m, m=m, colored=colored)
    ref = O.flowsys_lusgs_run(fs, fspr, tfspr, el, 6, 5, a0=0.2, turb_factor=1.7, kappa=1.1,
                              iterations=3)
    if where == "device":
        fsd, _, eld = on_device(fs, np.zeros(1), el)
        S = FS.FlowSysLUSGS(fsd, 6, 5, torch.from_numpy(fspr).cuda(),
                            torch.from_numpy(tfspr).cuda())
        S.set_element(0, eld)
    else:
        S = FS.FlowSysLUSGS(fs, 6, 5, fspr, tfspr)
        S.set_element(0, el)
    got = FS.run_lusgs(S, a0=0.2, turb_factor=1.7, kappa=1.1, iterations=3)
    S.close()
    assert sorted(got) == sorted(ref)
    for k, v in ref.items():
        g = np.asarray(got[k]).reshape(v.shape)
        assert np.array_equal(g, v), k


def test_lusgs_state_errors():
    from ucfd_sparse_amd import flowsys as FS
    from ucfd_sparse_amd import lib as L
    say("scalar LU-SGS error paths")
    mesh, flow, fs, fspr, tfspr, el = lusgs_case()
    S = FS.FlowSysLUSGS(fs, 6, 5, fspr, None)
    S.set_element(0, el)
    S.call("UCFDLUSGS_Pack", 0, 1.5, 0.3)
    with pytest.raises(L.UCFDError):
        S.call("UCFDLUSGS_RANSPrepare", 1.0)       # no turbulence spectral radius
    with pytest.raises(L.UCFDError):
        S.call("UCFDLUSGS_Pack", 3, 1.5, 0.3)      # no such element
    S.close()
    # two elements on one cell: reported, not a silent race
    eld = dict(el)
    eld["cell_ids"] = el["cell_ids"].copy()
    eld["cell_ids"][7] = eld["cell_ids"][3]
    S = FS.FlowSysLUSGS(fs, 6, 5, fspr, tfspr)
    S.set_element(0, eld)
    with pytest.raises(L.UCFDError):
        S.call("UCFDLUSGS_Pack", 0, 1.5, 0.3)
    S.close()
    B = FS.FlowSys(fs, 6, 5)                       # block LU-SGS handle
    with pytest.raises(L.UCFDError):
        B.call("UCFDLUSGS_NSPrepare", 1.0)
    B.close()


@pytest.mark.parametrize("colored", [False, True])
def test_lusgs_sweeps_without_prepare_and_other_kappa(colored):
    """The sweeps after a Prepare read the face records it wrote; sweeps
    with no Prepare before them read the caller's arrays.  Both equal the
    oracle bit for bit, also with a sweep kappa other than the Prepare's
    (kappa multiplies the spectral radius in the sweep itself, lusgs.c:196)."""
    from ucfd_sparse_amd import flowsys as FS
    mesh, flow, fs, fspr, tfspr, el = lusgs_case(seed=321, m=4, colored=colored)
    fsd, _, eld = on_device(fs, np.zeros(1), el)
    # none: no Prepare; both: NS then RANS Prepare (the RANS one reuses the
    # geometry records); rans: a RANS Prepare alone writes the geometry, the
    # NS sweeps read the caller's arrays
    for prepared in ("none", "both", "rans"):
        S = FS.FlowSysLUSGS(fsd, 6, 5, torch.from_numpy(fspr).cuda(),
                            torch.from_numpy(tfspr).cuda())
        S.set_element(0, eld)
        Or = O.FlowSysLUSGSOracle(fs, fspr, tfspr, el, 6, 5)
        S.call("UCFDLUSGS_Pack", 0, 1.3, 0.1)
        Or.pack(1.3, 0.1)
        if prepared == "both":
            S.call("UCFDLUSGS_NSPrepare", 0.9)
            Or.prepare(0.9)
        if prepared != "none":
            S.call("UCFDLUSGS_RANSPrepare", 0.9)
            Or.prepare(0.9, turb=True)
        for name, upper, turb in (("NSLower", False, False), ("NSUpper", True, False),
                                  ("RANSLower", False, True), ("RANSUpper", True, True)):
            S.call(f"UCFDLUSGS_{name}Sweep", 1.25)
            Or.sweep(upper, 1.25, turb=turb)
            assert np.array_equal(S.export(1), Or.du), (prepared, name)
        S.close()


def quad_mesh(m, n, seed):
    """2-D quads: the first four faces (x and y) of an m x n x 1 hex grid."""
    h = M.hex_mesh(m, n, 1, seed=seed)
    return M.Mesh(np.ascontiguousarray(h.nei_ele[:4]), np.ascontiguousarray(h.fnorm_vol[:4]),
                  np.ascontiguousarray(h.vec_fnorm[:4, :2]), h.orig_id, h.key)


@pytest.mark.parametrize("kind, nt, colored", [("quad", 1, False), ("quad", 2, True),
                                               ("tet", 2, True), ("tet", 4, False),
                                               ("hex", 1, False), ("hex", 2, True)])
def test_lusgs_two_dimensions_and_turbulence_rows(kind, nt, colored):
    """The scalar method in 2-D (nfvars = 4) and with 1 to 4 turbulence rows,
    on a random and on a colour-ordered numbering (the first schedule level
    then an index list or a range): every stage of two iterations bit for bit
    against the oracle (lusgs.c:8-321).  Hex cells (six faces) give a wave
    more positions than the Prepare stages in LDS, so the rest of them are
    written directly."""
    from ucfd_sparse_amd import flowsys as FS
    from ucfd_sparse_amd.layout import colored_layout
    seed = 61 + nt
    if kind == "quad":
        mesh = quad_mesh(9, 7, seed)
    elif kind == "hex":
        mesh = M.hex_mesh(7, 6, 5, seed=seed, permute=True)
    else:
        mesh = M.tet_mesh(3, seed=seed)
    nfv = mesh.ndims + 2
    flow = M.flow_fields(mesh, seed=seed, nvars=nfv + nt)
    if colored:
        mesh.key = None
        lay = colored_layout(mesh, flow, use_key=False)
        mesh, flow = lay.mesh, lay.flow
    fs = M.flowsys_system(mesh, flow, nv=nfv, seed=seed)
    rng = np.random.default_rng(seed + 5)
    fspr = rng.uniform(0.5, 2.0, fs.nfaces)
    tfspr = rng.uniform(0.2, 1.0, fs.nfaces)
    el = M.flowsys_elements(mesh, flow, nvars=nfv + nt, seed=seed)
    say(f"scalar LU-SGS {kind} ({mesh.ndims}-D), nt={nt}, colour-ordered={colored}: "
        f"{fs.nlocal} cells, {fs.nfaces} slots")
    ref = O.flowsys_lusgs_run(fs, fspr, tfspr, el, nfv + nt, nfv, a0=0.2, turb_factor=1.7,
                              kappa=1.1, iterations=2)
    fsd, _, eld = on_device(fs, np.zeros(1), el)
    S = FS.FlowSysLUSGS(fsd, nfv + nt, nfv, torch.from_numpy(fspr).cuda(),
                        torch.from_numpy(tfspr).cuda(), ndims=mesh.ndims)
    S.set_element(0, eld)
    got = FS.run_lusgs(S, a0=0.2, turb_factor=1.7, kappa=1.1, iterations=2)
    S.close()
    assert sorted(got) == sorted(ref)
    for k, v in ref.items():
        assert np.array_equal(np.asarray(got[k]).reshape(v.shape), v), k


# ------------------------------------------------------ two element types
def split_elements(el, seed):
    """The elements of one type dealt at random into two types, each with its
    own arrays ([nvars][neles_t] columns, per-element values) and cell_ids."""
    ne = len(el["cell_ids"])
    pick = np.random.default_rng(seed).permutation(ne)
    parts = [np.sort(pick[:ne // 3]), np.sort(pick[ne // 3:])]
    out = []
    for idx in parts:
        t = {}
        for k, v in el.items():
            v = np.asarray(v)
            t[k] = np.ascontiguousarray(v[..., idx] if v.ndim == 2 else v[idx])
        out.append(t)
    return parts, out


@pytest.mark.parametrize("colored", [False, True])
def test_block_method_two_element_types(colored):
    """Pack per element type writes only that type's cells (the record form
    skips cells no element of the type maps to): packing two types that
    share the cells equals packing them as one type, and every later stage
    (Prepare, sweeps, SubResidual and Update per type) matches the one-type
    oracle run bit for bit."""
    from ucfd_sparse_amd import flowsys as FS
    fs, tj, el = case(4, 73, colored, 6)
    ref = O.flowsys_run(fs, tj, el, 6, 5, kind="sa", iterations=2)
    parts, els = split_elements(el, 5)
    fsd, tjd, _ = on_device(fs, tj, el)
    S = FS.FlowSys(fsd, 6, 5, nelem=2, tjmat=tjd)
    for t, e in enumerate(els):
        S.set_element(t, {k: torch.from_numpy(v).cuda() for k, v in e.items()})
    for t in range(2):
        S.call("UCFDBLUSGS_Pack", t, 0.3)
        S.call("UCFDBLUSGS_SA_Pack", t, 1.5, 0.3)
    got = {"rank_rhs": S.export(FS.RHS), "diag_pack": S.export(FS.DIAG),
           "tdiag_pack": S.export(FS.TDIAG)}
    S.call("UCFDBLUSGS_NSPrepare")
    got["diag"] = S.export(FS.DIAG)
    S.call("UCFDBLUSGS_RANSPrepare")
    got["tdiag"] = S.export(FS.TDIAG)
    S.call("UCFDBLUSGS_Reset")
    for it in range(2):
        S.call("UCFDBLUSGS_NSLowerSweep")
        got[f"du_lower{it}"] = S.export(FS.DU)
        S.call("UCFDBLUSGS_NSUpperSweep")
        got[f"du_upper{it}"] = S.export(FS.DU)
        S.call("UCFDBLUSGS_RANSLowerSweep")
        S.call("UCFDBLUSGS_RANSUpperSweep")
        got[f"du_rans{it}"] = S.export(FS.DU)
        res = np.zeros_like(np.asarray(ref[f"resid{it}"]).reshape(6, -1))
        for t in range(2):
            S.call("UCFDBLUSGS_SubResidual", t)
            res[:, parts[t]] = S.element_array(t, "resid_out").cpu().numpy().reshape(6, -1)
        got[f"resid{it}"] = res
    up = np.zeros_like(np.asarray(ref["upts"]).reshape(6, -1))
    for t in range(2):
        S.call("UCFDBLUSGS_Update", t)
        up[:, parts[t]] = S.element_array(t, "uptsb").cpu().numpy().reshape(6, -1)
    got["upts"] = up
    S.close()
    say(f"two element types ({len(parts[0])} + {len(parts[1])} elements), colour-ordered={colored}")
    for k, v in ref.items():
        assert np.array_equal(np.asarray(got[k]).reshape(v.shape), v), k


@pytest.mark.parametrize("colored", [False, True])
def test_lusgs_two_element_types(colored):
    """The scalar method with the elements in two types: Pack and Update per
    type through the per-cell records, bit for bit equal to the one-type
    oracle run."""
    from ucfd_sparse_amd import flowsys as FS
    mesh, flow, fs, fspr, tfspr, el = lusgs_case(seed=306, m=3, colored=colored)
    ref = O.flowsys_lusgs_run(fs, fspr, tfspr, el, 6, 5, a0=0.2, turb_factor=1.7, kappa=1.1,
                              iterations=2)
    parts, els = split_elements(el, 9)
    fsd, _, _ = on_device(fs, np.zeros(1), el)
    S = FS.FlowSysLUSGS(fsd, 6, 5, torch.from_numpy(fspr).cuda(), torch.from_numpy(tfspr).cuda(),
                        nelem=2)
    for t, e in enumerate(els):
        S.set_element(t, {k: torch.from_numpy(v).cuda() for k, v in e.items()})
    for t in range(2):
        S.call("UCFDLUSGS_Pack", t, 1.7, 0.2)
    got = {"u": S.export(0), "du_pack": S.export(1), "diag_pack": S.export(2)}
    S.call("UCFDLUSGS_NSPrepare", 1.1)
    got["diag_ns"] = S.export(2)
    S.call("UCFDLUSGS_RANSPrepare", 1.1)
    got["diag"] = S.export(2)
    for it in range(2):
        for name, upper, key in (("NSLower", False, "du_nslower"), ("NSUpper", True, "du_nsupper"),
                                 ("RANSLower", False, "du_ranslower"),
                                 ("RANSUpper", True, "du_ransupper")):
            S.call(f"UCFDLUSGS_{name}Sweep", 1.1)
            got[f"{key}{it}"] = S.export(1)
    up = np.zeros_like(np.asarray(ref["upts"]).reshape(6, -1))
    for t in range(2):
        S.call("UCFDLUSGS_Update", t)
        up[:, parts[t]] = S.element_array(t, "uptsb").cpu().numpy().reshape(6, -1)
    got["upts"] = up
    S.close()
    say(f"scalar method, two element types, colour-ordered={colored}")
    assert sorted(got) == sorted(ref)
    for k, v in ref.items():
        assert np.array_equal(np.asarray(got[k]).reshape(v.shape), v), k
