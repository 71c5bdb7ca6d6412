"""Mesh files (CPU): the writePolymesh layout of src/LpmNetCDF_Impl.hpp:46-218
written by lpm_mesh_write_nc, read back by an independent netCDF reader
(scipy.io.netcdf_file) and by lpm_mesh_read_nc.

The topology restates FaceDivider<Geo,TriFace>::divide (src/LpmFaces.cpp:116-248)
and Edges::divide (src/LpmEdges.cpp:55-85).  It is pinned by the seed
tables at level 0 (the reference's mesh_seeds/icosTriSphereSeed.dat, kept in
tests/golden/) and, at every level, by what the construction fixes: each
face's edges join its consecutive vertices with the orientation that its
left/right entry says, edge children split their parent at the projected
midpoint (child 0 from the origin), face children are consecutive with
corner child k keeping corner k, leaves are counter-clockwise, and the leaf
complex is a closed surface of Euler characteristic 2."""
import os

import numpy as np
import pytest
import scipy.io

import lpmkokkos_amd as L

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "icosTriSphereSeed.dat")


def seed_tables():
    """Edges (orig, dest, left, right), face vertices and face edges of the seed file."""
    lines = [ln.split() for ln in open(GOLDEN)]
    i = next(k for k, ln in enumerate(lines) if ln and ln[0] == "edgeO")
    edges = np.array([[int(v) for v in ln] for ln in lines[i + 1:i + 31]])
    i = next(k for k, ln in enumerate(lines) if ln and ln[0] == "faceverts")
    fverts = np.array([[int(v) for v in ln] for ln in lines[i + 1:i + 21]])
    i = next(k for k, ln in enumerate(lines) if ln and ln[0] == "faceedges")
    fedges = np.array([[int(v) for v in ln] for ln in lines[i + 1:i + 21]])
    return edges, fverts, fedges


def read(path):
    f = scipy.io.netcdf_file(path, "r", mmap=False)
    v = {k: np.array(f.variables[k][:]) for k in f.variables}
    return f, v


@pytest.fixture(scope="module")
def files(tmp_path_factory):
    d = tmp_path_factory.mktemp("mesh")
    out = {}
    for lev in range(5):
        p = str(d / f"icostri_{lev}.nc")
        L.write_mesh_nc(lev, p)
        out[lev] = p
    return out


def test_level0_is_the_seed(files):
    f, v = read(files[0])
    edges, fverts, fedges = seed_tables()
    assert np.array_equal(v["edge_origs"], edges[:, 0]) and np.array_equal(v["edge_dests"], edges[:, 1])
    assert np.array_equal(v["edge_lefts"], edges[:, 2]) and np.array_equal(v["edge_rights"], edges[:, 3])
    assert np.array_equal(v["face_verts"], fverts) and np.array_equal(v["face_edges"], fedges)
    assert np.all(v["edge_parents"] == -1) and np.all(v["edge_kids"] == -1) and np.all(v["face_kids"] == -1)
    assert np.all(v["face_parents"] == -1) and np.all(v["face_tree_level"] == 0)


@pytest.mark.parametrize("lev", [0, 1, 3, 4])
def test_file_layout_and_arrays(files, lev):
    f, v = read(files[lev])
    m = L.TreeMesh(lev)
    assert f.version_byte == 2  # classic format, 64-bit offsets
    assert dict(f.dimensions) == {"ndim": 3, "nverts": m.nverts, "nedges": m.nedges, "nfaces": m.nfaces,
                                  "nfaceverts": 3, "two": 2, "four": 4}
    assert f._attributes["MeshSeed"] == b"IcosTriSphereSeed" and f._attributes["FaceKind"] == b"tri"
    assert f._attributes["Geom"] == b"SphereGeometry" and int(f._attributes["baseTreeDepth"]) == lev
    types = {k: f.variables[k].typecode() for k in f.variables}
    assert types["face_mask"] == "b" and types["face_area"] == "d" and types["edge_kids"] == "i"
    assert np.array_equal(v["phys_crds_verts"], m.vx) and np.array_equal(v["lag_crds_verts"], m.vlag)
    assert np.array_equal(v["phys_crds_faces"], m.fx) and np.array_equal(v["lag_crds_faces"], m.flag)
    assert np.array_equal(v["face_area"], m.farea) and np.array_equal(v["face_mask"].astype(np.uint8), m.fmask)
    assert np.array_equal(v["face_verts"], m.fverts) and np.array_equal(v["face_parents"], m.fparent)
    assert np.array_equal(v["face_centers"], np.arange(m.nfaces))


@pytest.mark.parametrize("lev", [2, 4])
def test_topology_invariants(files, lev):
    _, v = read(files[lev])
    o, d, lf, rt = v["edge_origs"], v["edge_dests"], v["edge_lefts"], v["edge_rights"]
    ek, ep = v["edge_kids"], v["edge_parents"]
    fv, fe, fk, fp, fl = v["face_verts"], v["face_edges"], v["face_kids"], v["face_parents"], v["face_tree_level"]
    vx = v["phys_crds_verts"]
    nf, ne = fv.shape[0], o.shape[0]
    # each face's edge k joins its vertices k, k+1, oriented as left/right says
    for f in range(nf):
        for k in range(3):
            e, a, b = fe[f, k], fv[f, k], fv[f, (k + 1) % 3]
            if lf[e] == f:
                assert (o[e], d[e]) == (a, b)
            else:
                assert rt[e] == f and (o[e], d[e]) == (b, a)
    # edge children: child 0 from the origin to the midpoint, child 1 on to the destination
    div = np.nonzero(ek[:, 0] >= 0)[0]
    assert np.all(ek[div, 1] == ek[div, 0] + 1)
    assert np.all(o[ek[div, 0]] == o[div]) and np.all(d[ek[div, 1]] == d[div])
    mid = d[ek[div, 0]]
    assert np.all(o[ek[div, 1]] == mid) and np.all(ep[ek[div, 0]] == div) and np.all(ep[ek[div, 1]] == div)
    mp = vx[o[div]] + vx[d[div]]
    mp /= np.linalg.norm(mp, axis=1)[:, None]
    assert np.abs(vx[mid] - mp).max() < 1e-15
    assert np.count_nonzero(ep >= 0) == 2 * len(div)
    # face children: consecutive, parent and level, corner child k keeps corner k,
    # the centre child has the parent's three edge midpoints
    fdiv = np.nonzero(fk[:, 0] >= 0)[0]
    assert np.array_equal(fdiv, np.nonzero(v["face_mask"])[0])
    assert np.all(v["face_area"][fdiv] == 0) and np.all(v["face_area"][fk[:, 0] < 0] > 0)
    for f in fdiv:
        kids = fk[f]
        assert np.array_equal(kids, kids[0] + np.arange(4))
        assert np.all(fp[kids] == f) and np.all(fl[kids] == fl[f] + 1)
        for k in range(3):
            assert fv[kids[k], k] == fv[f, k]
        mids = {d[ek[fe[f, k], 0]] for k in range(3)}
        assert set(fv[kids[3]]) == mids
    # leaves: counter-clockwise seen from outside, a closed surface
    leaf = fk[:, 0] < 0
    a, b, c = vx[fv[leaf, 0]], vx[fv[leaf, 1]], vx[fv[leaf, 2]]
    assert np.all(np.einsum("ij,ij->i", np.cross(b - a, c - a), a + b + c) > 0)
    leaf_edges = ek[:, 0] < 0
    assert vx.shape[0] - np.count_nonzero(leaf_edges) + np.count_nonzero(leaf) == 2
    uses = np.bincount(fe[leaf].ravel(), minlength=ne)
    assert np.all(uses[leaf_edges] == 2) and np.all(uses[~leaf_edges] == 0)
    # every level's complete edge set is kept: sum over l <= lev of nv_l + nf_l - 2 = 30 4^l
    assert ne == sum(30 * 4**k for k in range(lev + 1)) == L.MeshSeed().setMaxAllocations(lev)[1]


@pytest.mark.parametrize("lev", [0, 3])
def test_read_back(files, lev):
    m, r = L.TreeMesh(lev), L.TreeMesh.from_nc(files[lev])
    assert (r.nverts, r.nedges, r.nfaces, r.nleaves, r.level) == (m.nverts, m.nedges, m.nfaces, m.nleaves, lev)
    for k in ("vx", "vlag", "fx", "flag", "farea", "fmask", "fverts", "fparent"):
        assert np.array_equal(getattr(r, k), getattr(m, k)), k


def test_errors(tmp_path):
    bad = tmp_path / "bad.nc"
    bad.write_bytes(b"not a netcdf file at all")
    with pytest.raises(L.LpmError):
        L.TreeMesh.from_nc(str(bad))
    with pytest.raises(L.LpmError):
        L.TreeMesh.from_nc(str(tmp_path / "missing.nc"))
    with pytest.raises(L.LpmError):
        L.write_mesh_nc(13, str(tmp_path / "x.nc"))
    with pytest.raises(L.LpmError):
        L.write_mesh_nc(1, str(tmp_path / "no_such_dir" / "x.nc"))
    # a file scipy wrote without the mesh variables
    g = scipy.io.netcdf_file(str(tmp_path / "other.nc"), "w", version=2)
    g.createDimension("nverts", 3)
    g.createDimension("nfaces", 2)
    g.createDimension("ndim", 3)
    g.createVariable("phys_crds_verts", "d", ("nverts", "ndim"))[:] = 0.0
    g.close()
    with pytest.raises(L.LpmError):
        L.TreeMesh.from_nc(str(tmp_path / "other.nc"))


def test_reads_files_written_by_another_writer(files, tmp_path):
    """A mesh file written by scipy's netCDF writer (variables in another order,
    extra dimensions, variables and attributes of other types) reads back the
    same arrays: the reader follows the format, not this library's layout."""
    f, v = read(files[3])
    path = str(tmp_path / "other_writer.nc")
    g = scipy.io.netcdf_file(path, "w", version=2)
    g.title = "written by scipy"
    g.history = np.array([1.5, 2.5])
    g.createDimension("unused", 7)
    for name, length in f.dimensions.items():
        g.createDimension(name, length)
    g.createVariable("noise", "f", ("unused",))[:] = np.arange(7, dtype=np.float32)
    for name in reversed(sorted(f.variables)):
        src = f.variables[name]
        var = g.createVariable(name, src.typecode(), src.dimensions)
        var[:] = v[name]
        var.long_name = name.replace("_", " ")
    g.close()
    m, r = L.TreeMesh(3), L.TreeMesh.from_nc(path)
    assert (r.nverts, r.nedges, r.nfaces, r.nleaves) == (m.nverts, m.nedges, m.nfaces, m.nleaves)
    for k in ("vx", "vlag", "fx", "flag", "farea", "fmask", "fverts", "fparent"):
        assert np.array_equal(getattr(r, k), getattr(m, k)), k


def test_truncated_files_fail_cleanly(files, tmp_path):
    """Every truncation of a mesh file is reported as an error, never read as data."""
    data = open(files[1], "rb").read()
    for cut in (0, 3, 4, 40, 200, len(data) // 2, len(data) - 1):
        p = tmp_path / f"cut_{cut}.nc"
        p.write_bytes(data[:cut])
        with pytest.raises(L.LpmError):
            L.TreeMesh.from_nc(str(p))


@pytest.mark.parametrize("fmt", [0, 1, 2, 3])
def test_mesh_vtk(tmp_path, fmt):
    """PolyMesh2d::outputVtk: vertices, leaf faces and their areas (parsed here
    with the GPU tests' independent VTK readers)."""
    from test_gpu_files import read_legacy, read_xml
    m = L.TreeMesh(3)
    path = str(tmp_path / ("m.vtp" if fmt >= 2 else "m.vtk"))
    L.write_mesh_vtk(3, path, fmt)
    d = read_xml(path) if fmt >= 2 else read_legacy(path)
    leaf = m.fmask == 0
    assert np.array_equal(d["points"], m.vx)
    polys = d["polys"] if fmt >= 2 else d["polys"][:, 1:]
    assert np.array_equal(polys, m.fverts[leaf])
    assert list(d["point"]) == [] and list(d["cell"]) == ["area"]
    assert np.array_equal(d["cell"]["area"].ravel(), m.farea[leaf])
    with pytest.raises(L.LpmError):
        L.write_mesh_vtk(3, path, 7)


def test_cubed_sphere_files(tmp_path):
    """The cubed-sphere tree (MeshSeed<CubedSphereSeed>, src/LpmFaces.cpp:251-381)
    in the writePolymesh layout: four vertices per face, MeshSeed
    "CubedSphereSeed", FaceKind "quad"; the arrays equal the oracle's
    restatement bit for bit, level 0 equals the seed file's tables, and the
    file reads back; the VTK mesh holds quadrilaterals."""
    from oracle import oracle as O
    from test_gpu_files import read_legacy, read_xml
    seed = L.MeshSeed(L.CubedSphereSeed)
    crds, sedges, sfv, sfe = O.read_cubed_seed()
    for lev in (0, 2, 4):
        m = L.TreeMesh(lev, seed)
        o = O.Mesh(lev, kind="cubed")
        for k in ("vx", "vlag", "fx", "flag", "farea", "fmask", "fverts", "fparent"):
            assert np.array_equal(getattr(m, k), getattr(o, k)), (lev, k)
        p = str(tmp_path / f"cubedsphere_{lev}.nc")
        L.write_mesh_nc(lev, p, seed)
        f, v = read(p)
        assert dict(f.dimensions) == {"ndim": 3, "nverts": m.nverts, "nedges": m.nedges, "nfaces": m.nfaces,
                                      "nfaceverts": 4, "two": 2, "four": 4}
        assert f._attributes["MeshSeed"] == b"CubedSphereSeed" and f._attributes["FaceKind"] == b"quad"
        assert np.array_equal(v["face_verts"], m.fverts) and np.array_equal(v["face_area"], m.farea)
        if lev == 0:
            assert np.array_equal(v["edge_origs"], sedges[:, 0]) and np.array_equal(v["edge_rights"], sedges[:, 3])
            assert np.array_equal(v["face_verts"], sfv) and np.array_equal(v["face_edges"], sfe)
            assert np.array_equal(m.vx, crds[:8]) and np.array_equal(m.fx, crds[8:])
        r = L.TreeMesh.from_nc(p)
        assert r.seed_type is L.CubedSphereSeed and r.level == lev
        for k in ("vx", "fx", "farea", "fmask", "fverts", "fparent"):
            assert np.array_equal(getattr(r, k), getattr(m, k)), k
        # each face's edge k joins its vertices k and k+1 as left/right says
        o_, d_, lf = v["edge_origs"], v["edge_dests"], v["edge_lefts"]
        for fi in range(m.nfaces):
            for k in range(4):
                e, a, b = v["face_edges"][fi, k], m.fverts[fi, k], m.fverts[fi, (k + 1) % 4]
                assert (o_[e], d_[e]) == ((a, b) if lf[e] == fi else (b, a))
    leaf = m.fmask == 0
    for fmt in (0, 3):
        path = str(tmp_path / ("c.vtp" if fmt >= 2 else "c.vtk"))
        L.write_mesh_vtk(4, path, fmt, seed)
        d = read_xml(path) if fmt >= 2 else read_legacy(path)
        polys = d["polys"] if fmt >= 2 else d["polys"][:, 1:]
        assert polys.shape[1] == 4 and np.array_equal(polys, m.fverts[leaf])
        assert np.array_equal(d["cell"]["area"].ravel(), m.farea[leaf])
