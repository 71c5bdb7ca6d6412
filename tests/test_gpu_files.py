"""Checkpoint / resume and VTK output of a resident state.

* A run checkpointed after k steps and resumed in a fresh context takes the
  same next steps, bit for bit, as the run that never stopped (every scheme
  and precision, states built by level or from host arrays, before or after
  the first step).
* The checkpoint is a netCDF file with the writePolymesh variables
  (src/LpmNetCDF_Impl.hpp:46-218) and the BVESphere fields under their view
  labels (src/LpmBVESphere.cpp:13-20), read here by scipy's reader.
* VTK files hold the vertices, the leaf faces and the fields with the names of
  BVESphere::outputVtk (legacy) and BVESphere::addFieldsToVtk (XML), read here
  by small independent parsers."""
import base64
import math
import xml.etree.ElementTree as ET

import numpy as np
import pytest
import scipy.io

import lpmkokkos_amd as L

pytestmark = pytest.mark.gpu

TWO_PI = 2 * math.pi
DT = 0.005


def field():
    ic = L.VorticityInitialCondition(sbr=0.5, rh_n=4, rh_w=0.1, rh_K=0.1)
    return ic.with_vortices(5, 12, 0.3, 0.15)


def states_equal(a, b):
    for k in ("vx", "vzeta", "vu", "fx", "fzeta", "fu"):
        assert np.array_equal(a[k], b[k]), k


def level_ctx(level):
    c = L.Context(0)
    c.set_state_level(level, field(), subtract_mean=True)
    return c


def host_ctx(level):
    m = L.TreeMesh(level)
    ic = field()
    c = L.Context(0)
    c.set_state(m.vx, ic.eval_many(m.vx), m.fx, ic.eval_many(m.fx), m.farea, m.fmask)
    return c


@pytest.mark.parametrize("scheme,prec", [(L.SCHEME_LAGR, L.PREC_FP32C), (L.SCHEME_LAGR, L.PREC_FP64),
                                         (L.SCHEME_REF, L.PREC_FP64)])
@pytest.mark.parametrize("build", ["level", "host"])
def test_resume_is_bitwise(tmp_path, scheme, prec, build):
    level = 4
    a = level_ctx(level) if build == "level" else host_ctx(level)
    a.advance(DT, TWO_PI, scheme, prec, nsteps=3)
    path = str(tmp_path / "ck.nc")
    a.checkpoint(path, time=3 * DT, omega=TWO_PI)
    a.advance(DT, TWO_PI, scheme, prec, nsteps=3)
    b = L.Context(0)
    t, om = b.restore(path)
    assert (t, om) == (3 * DT, TWO_PI)
    b.advance(DT, TWO_PI, scheme, prec, nsteps=3)
    states_equal(a.get_state(), b.get_state())
    a.close()
    b.close()


def test_resume_before_first_step_keeps_psi(tmp_path):
    a = level_ctx(3)
    a.solve(L.PREC_FP64, L.SCHEME_LAGR, want_psi=True)
    path = str(tmp_path / "ck0.nc")
    a.checkpoint(path, omega=TWO_PI)
    b = L.Context(0)
    b.restore(path)
    sa, sb = a.get_state(), b.get_state()
    states_equal(sa, sb)
    assert np.array_equal(sa["vpsi"], sb["vpsi"]) and np.array_equal(sa["fpsi"], sb["fpsi"])
    assert np.abs(sa["vpsi"]).max() > 0
    for c in (a, b):
        c.advance(DT, TWO_PI, L.SCHEME_LAGR, L.PREC_FP32C, nsteps=2)
    states_equal(a.get_state(), b.get_state())
    a.close()
    b.close()


def test_checkpoint_file_contents(tmp_path):
    level = 3
    m = L.TreeMesh(level)
    a = level_ctx(level)
    s0 = a.get_state()
    a.advance(DT, TWO_PI, L.SCHEME_LAGR, L.PREC_FP64, nsteps=2)
    path = str(tmp_path / "ck.nc")
    a.checkpoint(path, time=2 * DT, omega=TWO_PI)
    s = a.get_state()
    f = scipy.io.netcdf_file(path, "r", mmap=False)
    v = {k: np.array(f.variables[k][:]) for k in f.variables}
    assert f.dimensions["nverts"] == m.nverts and f.dimensions["nedges"] == m.nedges
    assert f.dimensions["nfaces"] == m.nfaces and int(f._attributes["baseTreeDepth"]) == level
    assert float(f._attributes["time"]) == 2 * DT and float(f._attributes["Omega"]) == TWO_PI
    # current positions; Lagrangian coordinates and topology of the tree
    assert np.array_equal(v["phys_crds_verts"], s["vx"]) and np.array_equal(v["phys_crds_faces"], s["fx"])
    assert np.array_equal(v["lag_crds_verts"], m.vlag) and np.array_equal(v["face_verts"], m.fverts)
    assert np.array_equal(v["face_area"], m.farea) and np.array_equal(v["face_mask"].astype(np.uint8), m.fmask)
    # fields under the reference's labels
    assert np.array_equal(v["relVortVerts"], s["vzeta"]) and np.array_equal(v["velocityVerts"], s["vu"])
    lv = m.fmask == 0
    assert np.array_equal(v["relVortFaces"], s["fzeta"]) and np.array_equal(v["velocityFaces"][lv], s["fu"][lv])
    assert f.variables["relVortVerts"].units == b""
    # absolute vorticity: that of the initial state, conserved
    av0 = s0["vzeta"] + 2 * TWO_PI * s0["vx"][:, 2]
    assert np.abs(v["absVortVerts"] - av0).max() < 1e-14 * np.abs(av0).max()
    af0 = s0["fzeta"][lv] + 2 * TWO_PI * s0["fx"][lv, 2]
    assert np.abs(v["absVortFaces"][lv] - af0).max() < 1e-14 * np.abs(af0).max()
    assert np.abs(v["absVortVerts"] - 2 * TWO_PI * s["vx"][:, 2] - s["vzeta"]).max() < 1e-13
    assert not int(f._attributes["streamfn_valid"]) and np.all(v["streamFnVerts"] == 0)
    a.close()


def test_checkpoint_without_tree(tmp_path):
    """A state set from arrays, checkpointed with no level: particles only."""
    a = host_ctx(2)
    path = str(tmp_path / "p.nc")
    a.checkpoint(path, omega=TWO_PI)
    f = scipy.io.netcdf_file(path, "r", mmap=False)
    assert "nedges" not in f.dimensions and "face_verts" not in f.variables
    assert "relVortFaces" in f.variables
    with pytest.raises(L.LpmError):
        a.write_vtk(str(tmp_path / "x.vtk"), TWO_PI)  # no connectivity
    a.checkpoint(str(tmp_path / "q.nc"), omega=TWO_PI, level=2)  # the caller names the tree
    f = scipy.io.netcdf_file(str(tmp_path / "q.nc"), "r", mmap=False)
    assert f.dimensions["nedges"] == L.TreeMesh(2).nedges
    with pytest.raises(L.LpmError):
        a.checkpoint(str(tmp_path / "r.nc"), omega=TWO_PI, level=3)  # sizes do not match
    with pytest.raises(L.LpmError):
        a.checkpoint(str(tmp_path / "no_dir" / "r.nc"), omega=TWO_PI)
    b = L.Context(0)
    with pytest.raises(L.LpmError):
        b.restore(str(tmp_path / "missing.nc"))
    # a state without vertices: netCDF classic has no fixed dimension of length 0
    m = L.TreeMesh(2)
    b.set_state(np.zeros((0, 3)), np.zeros(0), m.fx, np.ones(m.nfaces), m.farea, m.fmask)
    with pytest.raises(L.LpmError):
        b.checkpoint(str(tmp_path / "z.nc"))
    a.close()
    b.close()


def test_checkpoint_keeps_the_state_areas(tmp_path):
    """A state whose areas differ from the tree's (a caller's mesh) is
    checkpointed with the tree level named: the file holds the state's areas,
    and the resumed run is bit-identical to the one that never stopped.  A
    mask that differs from the tree's is refused with a tree level."""
    level = 3
    m = L.TreeMesh(level)
    rng = np.random.default_rng(11)
    area = m.farea * (1 + 1e-3 * rng.uniform(-1, 1, m.nfaces))
    ic = field()
    a = L.Context(0)
    a.set_state(m.vx, ic.eval_many(m.vx), m.fx, ic.eval_many(m.fx), area, m.fmask)
    a.advance(DT, TWO_PI, L.SCHEME_LAGR, L.PREC_FP64, nsteps=2)
    path = str(tmp_path / "areas.nc")
    a.checkpoint(path, time=2 * DT, omega=TWO_PI, level=level)
    f = scipy.io.netcdf_file(path, "r", mmap=False)
    lv = m.fmask == 0
    assert np.array_equal(np.array(f.variables["face_area"][:])[lv], area[lv])
    a.advance(DT, TWO_PI, L.SCHEME_LAGR, L.PREC_FP64, nsteps=2)
    b = L.Context(0)
    b.restore(path)
    b.advance(DT, TWO_PI, L.SCHEME_LAGR, L.PREC_FP64, nsteps=2)
    states_equal(a.get_state(), b.get_state())
    mask = m.fmask.copy()
    mask[np.nonzero(lv)[0][:4]] = 1  # four leaves masked: not the tree's mask
    c = L.Context(0)
    c.set_state(m.vx, ic.eval_many(m.vx), m.fx, ic.eval_many(m.fx), area, mask)
    with pytest.raises(L.LpmError):
        c.checkpoint(str(tmp_path / "mask.nc"), omega=TWO_PI, level=level)
    c.checkpoint(str(tmp_path / "mask.nc"), omega=TWO_PI)  # without the tree: particles only
    for x in (a, b, c):
        x.close()


# --- VTK readers -------------------------------------------------------------

class _Cursor:
    def __init__(self, data):
        self.b, self.p = data, 0

    def line(self):
        e = self.b.index(b"\n", self.p)
        s = self.b[self.p:e].decode()
        self.p = e + 1
        return s

    def values(self, n, binary, dtype):
        if binary:
            size = np.dtype(dtype).itemsize
            a = np.frombuffer(self.b[self.p:self.p + n * size], dtype=">" + np.dtype(dtype).str[1:])
            self.p += n * size
            assert self.b[self.p:self.p + 1] == b"\n"
            self.p += 1
            return a.astype(dtype)
        out = []
        while len(out) < n:
            out += self.line().split()
        return np.array(out, dtype=dtype)


def legacy_decode(name):
    """VTK legacy array names carry blanks and other bytes as %XX."""
    import re
    return re.sub(r"%([0-9A-F]{2})", lambda m: chr(int(m.group(1), 16)), name)


def read_legacy(path):
    c = _Cursor(open(path, "rb").read())
    assert c.line().startswith("# vtk DataFile")
    c.line()
    binary = c.line() == "BINARY"
    assert c.line() == "DATASET POLYDATA"
    _, n, _ = c.line().split()
    n = int(n)
    pts = c.values(3 * n, binary, np.float64).reshape(n, 3)
    tag, nc, size = c.line().split()
    assert tag == "POLYGONS"
    nsize = int(size) // int(nc)  # every polygon has the same vertex count here (3 or 4, plus the count)
    conn = c.values(int(size), binary, np.int32).reshape(int(nc), nsize)
    assert np.all(conn[:, 0] == nsize - 1)
    out = {"points": pts, "polys": conn, "point": {}, "cell": {}}
    while c.p < len(c.b):
        kind, count = c.line().split()
        key = "point" if kind == "POINT_DATA" else "cell"
        _, _, narr = c.line().split()
        for _ in range(int(narr)):
            name, ncomp, nt, _ = c.line().split()
            out[key][legacy_decode(name)] = c.values(int(ncomp) * int(nt), binary, np.float64).reshape(int(nt), int(ncomp))
    return out


def read_xml(path):
    root = ET.parse(path).getroot()
    assert root.get("type") == "PolyData" and root.get("byte_order") == "LittleEndian"

    def arr(el):
        dt = np.float64 if el.get("type") == "Float64" else np.int64
        if el.get("format") == "binary":
            raw = base64.b64decode(el.text.strip())
            nbytes = int(np.frombuffer(raw[:8], "<u8")[0])
            a = np.frombuffer(raw[8:8 + nbytes], dt)
            assert len(raw) == 8 + nbytes
        else:
            a = np.array(el.text.split(), dtype=dt)
        return a.reshape(-1, int(el.get("NumberOfComponents", "1")))

    piece = root.find("PolyData/Piece")
    out = {"points": arr(piece.find("Points/DataArray")), "point": {}, "cell": {}}
    for key, tag in (("point", "PointData"), ("cell", "CellData")):
        for el in piece.find(tag).findall("DataArray"):
            out[key][el.get("Name")] = arr(el)
    polys = {el.get("Name"): arr(el).ravel() for el in piece.find("Polys").findall("DataArray")}
    nfv = int(polys["offsets"][0])  # every polygon has the same vertex count here (3 or 4)
    assert np.array_equal(polys["offsets"], nfv * np.arange(1, len(polys["offsets"]) + 1))
    out["polys"] = polys["connectivity"].reshape(-1, nfv)
    assert int(piece.get("NumberOfPoints")) == len(out["points"]) and int(piece.get("NumberOfPolys")) == len(out["polys"])
    return out


@pytest.mark.parametrize("fmt", [L.VTK_LEGACY_ASCII, L.VTK_LEGACY_BINARY, L.VTK_XML_ASCII, L.VTK_XML_BINARY])
def test_vtk_contents(tmp_path, fmt):
    level = 3
    m = L.TreeMesh(level)
    a = level_ctx(level)
    a.advance(DT, TWO_PI, L.SCHEME_LAGR, L.PREC_FP64, nsteps=1)
    a.solve(L.PREC_FP64, L.SCHEME_LAGR, want_psi=True)
    s = a.get_state()
    path = str(tmp_path / ("s.vtp" if fmt >= L.VTK_XML_ASCII else "s.vtk"))
    a.write_vtk(path, TWO_PI, fmt=fmt)
    d = read_xml(path) if fmt >= L.VTK_XML_ASCII else read_legacy(path)
    leaf = m.fmask == 0
    assert np.array_equal(d["points"], s["vx"])
    polys = d["polys"] if fmt >= L.VTK_XML_ASCII else d["polys"][:, 1:]
    assert fmt >= L.VTK_XML_ASCII or np.all(d["polys"][:, 0] == 3)
    assert np.array_equal(polys, m.fverts[leaf])
    if fmt >= L.VTK_XML_ASCII:
        names = ("relvort", "absvort", "stream_fn", "velocity")
    else:
        names = ("relVort", "absVort", "streamFn", "velocity")
    assert list(d["point"]) == list(names) and list(d["cell"]) == ["area"] + list(names)
    zeta, absv, psi, vel = names
    assert np.array_equal(d["point"][zeta].ravel(), s["vzeta"]) and np.array_equal(d["point"][psi].ravel(), s["vpsi"])
    assert np.array_equal(d["point"][vel], s["vu"]) and np.abs(s["vpsi"]).max() > 0
    assert np.array_equal(d["cell"]["area"].ravel(), m.farea[leaf])
    assert np.array_equal(d["cell"][zeta].ravel(), s["fzeta"][leaf]) and np.array_equal(d["cell"][vel], s["fu"][leaf])
    assert np.array_equal(d["cell"][psi].ravel(), s["fpsi"][leaf])
    assert np.abs(d["point"][absv].ravel() - s["vzeta"] - 2 * TWO_PI * s["vx"][:, 2]).max() < 1e-13
    a.close()


def test_sphere_mirror_output_and_restart(tmp_path):
    """BVESphere::outputVtk and a checkpoint / restore of the mirror classes."""
    lev = 3
    nv, ne, nf = L.MeshSeed().setMaxAllocations(lev)
    sphere = L.BVESphere(nv, ne, nf)
    sphere.treeInit(lev, L.MeshSeed())
    sphere.set_omega(TWO_PI)
    sphere.init_vorticity(field())
    rk = L.BVERK4(DT, TWO_PI)
    rk.init(nv, nf)
    args = (sphere.physVerts.crds, sphere.relVortVerts, sphere.velocityVerts, sphere.physFaces.crds,
            sphere.relVortFaces, sphere.velocityFaces, sphere.faces.area, sphere.faces.mask)
    rk.advance_timestep(*args, nsteps=2)
    sphere.outputVtk(str(tmp_path / "bve.vtk"))
    sphere.outputVtk(str(tmp_path / "bve.vtp"))
    assert len(read_legacy(str(tmp_path / "bve.vtk"))["polys"]) == sphere.mesh.nleaves
    assert len(read_xml(str(tmp_path / "bve.vtp"))["polys"]) == sphere.mesh.nleaves
    sphere.checkpoint(str(tmp_path / "bve.nc"))
    other = L.BVESphere(nv, ne, nf)
    other.restore(str(tmp_path / "bve.nc"))
    assert other.t == sphere.t and other.Omega == TWO_PI and other.mesh.level == lev
    rk.advance_timestep(*args, nsteps=2)
    args2 = (other.physVerts.crds, other.relVortVerts, other.velocityVerts, other.physFaces.crds,
             other.relVortFaces, other.velocityFaces, other.faces.area, other.faces.mask)
    rk.advance_timestep(*args2, nsteps=2)
    states_equal(sphere.ctx.get_state(), other.ctx.get_state())


def test_vtk_caller_arrays(tmp_path):
    """Fields added after the state's own, as the reference's driver adds its
    error arrays: vertex arrays whole, face arrays restricted to the leaves."""
    level = 2
    m = L.TreeMesh(level)
    a = level_ctx(level)
    a.solve(L.PREC_FP64, L.SCHEME_LAGR)
    rng = np.random.default_rng(3)
    pv, pf, sv = rng.normal(size=(m.nverts, 3)), rng.normal(size=(m.nfaces, 3)), rng.normal(size=m.nverts)
    arrays = [("vertex_velocity_error", False, pv), ("face_velocity_error", True, pf), ("tracer", False, sv)]
    leaf = m.fmask == 0
    for fmt in (L.VTK_LEGACY_BINARY, L.VTK_XML_ASCII):
        path = str(tmp_path / ("x.vtp" if fmt == L.VTK_XML_ASCII else "x.vtk"))
        a.write_vtk(path, TWO_PI, fmt=fmt, arrays=arrays)
        d = read_xml(path) if fmt == L.VTK_XML_ASCII else read_legacy(path)
        assert list(d["point"])[4:] == ["vertex_velocity_error", "tracer"]
        assert list(d["cell"])[5:] == ["face_velocity_error"]
        assert np.array_equal(d["point"]["vertex_velocity_error"], pv)
        assert np.array_equal(d["point"]["tracer"].ravel(), sv)
        assert np.array_equal(d["cell"]["face_velocity_error"], pf[leaf])
    with pytest.raises(L.LpmError):
        a.write_vtk(str(tmp_path / "y.vtk"), TWO_PI, arrays=[("", False, sv)])
    with pytest.raises(ValueError):
        a.write_vtk(str(tmp_path / "y.vtk"), TWO_PI, arrays=[("short", True, sv)])
    a.close()


def test_cpp_driver_writes_the_reference_files(tmp_path):
    """bve_driver -o: XML files named as the reference driver names them, with
    the solid-body error arrays (zero velocity error at t = 0 up to the
    quadrature error, zero position error at t = 0)."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "lpmkokkos_amd", "bve_driver")
    case = str(tmp_path / "bve_test")
    r = subprocess.run([exe, "-d", "3", "-dt", "0.01", "-tf", "0.04", "-f", "2", "-o", case, "-scheme", "lagr",
                        "-seed", "icos"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    names = sorted(p.name for p in tmp_path.iterdir())
    assert names == ["bve_testtri3_dt0.01_0000.vtp", "bve_testtri3_dt0.01_0002.vtp", "bve_testtri3_dt0.01_0004.vtp"]
    d0 = read_xml(str(tmp_path / names[0]))
    assert list(d0["point"]) == ["relvort", "absvort", "stream_fn", "velocity", "u dot x", "vertex_velocity_error",
                                 "vertex_position_error"]
    assert list(d0["cell"]) == ["area", "relvort", "absvort", "stream_fn", "velocity", "u dot x",
                                "face_velocity_error", "face_position_error"]
    assert np.abs(d0["point"]["u dot x"]).max() < 1e-12 and np.abs(d0["cell"]["u dot x"]).max() < 1e-12
    assert np.all(d0["point"]["vertex_position_error"] == 0) and np.abs(d0["cell"]["face_velocity_error"]).max() < 0.1
    d2 = read_xml(str(tmp_path / names[2]))
    assert 0 < np.abs(d2["cell"]["face_position_error"]).max() < 0.01
    # the stream function of each output step is solved at that step (the
    # reference's post-timestep solve): for solid-body rotation about z it is a
    # function of z, which each particle keeps, so every output has the first's
    p0, p2 = d0["point"]["stream_fn"].ravel(), d2["point"]["stream_fn"].ravel()
    assert np.abs(p2).max() > 0.1 and np.abs(p2 - p0).max() < 1e-3 * np.abs(p0).max()
    # the driver's default seed, the cubed sphere: "quad" files of quadrilaterals
    qdir = tmp_path / "quad"
    qdir.mkdir()
    r = subprocess.run([exe, "-d", "3", "-dt", "0.01", "-tf", "0.02", "-f", "2", "-o", str(qdir / "bve_test")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    qnames = sorted(p.name for p in qdir.iterdir())
    assert qnames == ["bve_testquad3_dt0.01_0000.vtp", "bve_testquad3_dt0.01_0002.vtp"]
    q = read_xml(str(qdir / qnames[0]))
    m = L.TreeMesh(3, L.MeshSeed(L.CubedSphereSeed))
    assert np.array_equal(q["polys"], m.fverts[m.fmask == 0]) and len(q["points"]) == m.nverts


def test_sphere_tracers_in_vtk(tmp_path):
    """BVESphere tracers (create_tracer, and the nq built-in ones) are written
    after the state's fields, leaves only for faces, under their names -- blanks
    and markup encoded as VTK encodes them (%XX in legacy files, entities in XML)."""
    lev = 2
    nv, ne, nf = L.MeshSeed().setMaxAllocations(lev)
    sphere = L.BVESphere(nv, ne, nf, nq=1)
    sphere.treeInit(lev, L.MeshSeed())
    sphere.init_vorticity(field())
    k = sphere.create_tracer('u dot x <"%">')
    assert sphere.tracer_names == ["tracer0", 'u dot x <"%">'] and k == 1
    sphere.tracer_verts[0][:nv] = np.arange(nv)
    sphere.tracer_faces[1][:nf] = -np.arange(nf)
    leaf = sphere.mesh.fmask == 0
    for name in ("t.vtk", "t.vtp"):
        sphere.outputVtk(str(tmp_path / name))
        raw = open(tmp_path / name, "rb").read()
        d = read_xml(str(tmp_path / name)) if name.endswith(".vtp") else read_legacy(str(tmp_path / name))
        assert list(d["point"])[4:] == ["tracer0", 'u dot x <"%">']
        assert list(d["cell"])[5:] == ["tracer0", 'u dot x <"%">']
        assert np.array_equal(d["point"]["tracer0"].ravel(), np.arange(nv))
        assert np.array_equal(d["cell"]['u dot x <"%">'].ravel(), -np.arange(nf)[leaf])
        assert (b"u%20dot%20x%20<\"%25\">" in raw) if name.endswith(".vtk") else (b"u dot x &lt;&quot;%&quot;&gt;" in raw)

