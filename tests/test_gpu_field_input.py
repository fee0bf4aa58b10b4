"""PIC field input pipeline on the GPU (tp_gpu_load_field_grid and its
callers: tp_gpu_set_field_grid_planes, tp_host_load_pic_gda, tp_h5_load_pic).

The raw component planes go to the GPU in slabs through pinned staging
memory and a kernel scales, transposes and interleaves them into the nodes.
The nodes read back (tp_gpu_get_field_nodes) must equal, bit for bit, the
reference's loaders' arrays in closed form -- (float)(v * scale) in double,
stored x fastest -- and the nodes tp_gpu_set_field_grid(_grid2) builds from
those arrays; the gathers then evaluate the same field.  Covered: both plane
orders at sizes that leave partial LDS tiles, slabs of one row and ragged
last slabs, two time slices, the reference's .gda and HDF5 files (the HDF5
golden vectors of the reference's own loaders), a reader that fails, and a
256 x 192 x 160 grid.
"""
import os

import numpy as np
import pytest

from conftest import golden
from test_particle_amd import abi

pytestmark = pytest.mark.gpu


def case(msg):
    print(f"[case] {msg}", flush=True)


@pytest.fixture(scope="module")
def gpu():
    from test_particle_amd.gpu import Tracker
    t = Tracker(0)
    yield t
    t.close()


def expected_nodes(planes, order, e0, b0):
    """(nz, ny, nx, 8) nodes of raw planes, as the reference's loaders scale
    and order them (src/electric_field.c:276, :475-503)."""
    p = planes.astype(np.float64)
    if order == abi.TP_PLANES_X_SLOWEST:
        p = p.transpose(0, 3, 2, 1)                   # (c, x, y, z) -> (c, z, y, x)
    out = np.zeros((*p.shape[1:], 8), np.float32)
    for c in range(6):
        out[..., c] = (p[c] * (e0 if c < 3 else b0)).astype(np.float32)
    return out


def aos(nodes):
    """float3 E and B arrays (nz, ny, nx, 3) of the nodes."""
    return np.ascontiguousarray(nodes[..., :3]), np.ascontiguousarray(nodes[..., 3:6])


def dims(planes, order):
    a, b, c = planes.shape[1:]
    return (c, b, a) if order == abi.TP_PLANES_X_FASTEST else (a, b, c)   # (nx, ny, nz)


def rand_points(rng, lo, hi, n, t=0.0):
    p = np.column_stack([rng.uniform(lo[k], hi[k], n) for k in range(3)] + [np.full(n, t)])
    return p


@pytest.mark.parametrize("order", [abi.TP_PLANES_X_FASTEST, abi.TP_PLANES_X_SLOWEST])
@pytest.mark.parametrize("shape", [(13, 7, 11), (37, 5, 70), (2, 2, 2), (64, 3, 33)])
def test_planes_equal_reference_layout(gpu, order, shape):
    nx, ny, nz = shape
    case(f"planes order={order} shape={shape}")
    rng = np.random.default_rng(nx * 1000 + nz)
    pshape = (6, nz, ny, nx) if order == abi.TP_PLANES_X_FASTEST else (6, nx, ny, nz)
    planes = rng.normal(0, 1, pshape).astype(np.float32)
    lo, hi = (0.0, -1.0, 2.0), (5.0, 3.0, 5.5)
    e0, b0 = 0.37, 2.9
    gpu.set_field_grid_planes(planes, lo, hi, order=order, e0=e0, b0=b0)
    got = gpu.field_nodes(shape)
    want = expected_nodes(planes, order, e0, b0)
    assert np.array_equal(got, want)
    # the same nodes and field as tp_gpu_set_field_grid of the reference's arrays
    pts = rand_points(rng, lo, hi, 500)
    f_planes = gpu.eval_field(pts)
    e, b = aos(want)
    gpu.set_field_grid(e, b, lo, hi)
    assert np.array_equal(gpu.field_nodes(shape), want)
    assert np.array_equal(gpu.eval_field(pts), f_planes)


@pytest.mark.parametrize("order", [abi.TP_PLANES_X_FASTEST, abi.TP_PLANES_X_SLOWEST])
@pytest.mark.parametrize("mode", [abi.TP_GATHER_REF_COMPAT, abi.TP_GATHER_TEXTBOOK])
def test_two_slices(gpu, order, mode):
    nx, ny, nz = 19, 9, 14
    case(f"two slices order={order} mode={mode}")
    rng = np.random.default_rng(5 + order + 2 * mode)
    pshape = (6, nz, ny, nx) if order == abi.TP_PLANES_X_FASTEST else (6, nx, ny, nz)
    pn = rng.normal(0, 1, pshape).astype(np.float32)
    pp = rng.normal(0, 1, pshape).astype(np.float32)
    lo, hi = (-1.0, 0.0, 0.5), (3.0, 2.0, 4.0)
    gpu.set_field_grid_planes(pn, lo, hi, order=order, e0=0.5, b0=1.5, planes_prev=pp,
                              grid_dt=0.4, mode=mode)
    wn, wp = expected_nodes(pn, order, 0.5, 1.5), expected_nodes(pp, order, 0.5, 1.5)
    assert np.array_equal(gpu.field_nodes((nx, ny, nz), 0), wn)
    assert np.array_equal(gpu.field_nodes((nx, ny, nz), 1), wp)
    pts = rand_points(rng, lo, hi, 400, t=0.13)
    f_planes = gpu.eval_field(pts)
    en, bn = aos(wn)
    ep, bp = aos(wp)
    gpu.set_field_grid2(en, bn, ep, bp, lo, hi, 0.4, mode=mode)
    assert np.array_equal(gpu.eval_field(pts), f_planes)


@pytest.mark.parametrize("order", [abi.TP_PLANES_X_FASTEST, abi.TP_PLANES_X_SLOWEST])
@pytest.mark.parametrize("rows", [1, 3])
def test_small_slabs_cover_the_grid(gpu, order, rows):
    """Slabs of one row and of three rows (a ragged last slab): every slab is
    asked for once, in order, and the nodes are whole."""
    nx, ny, nz = 11, 6, 10
    case(f"slabs order={order} rows={rows}")
    rng = np.random.default_rng(40 + rows)
    pshape = (6, nz, ny, nx) if order == abi.TP_PLANES_X_FASTEST else (6, nx, ny, nz)
    pl = [rng.normal(0, 1, pshape).astype(np.float32) for _ in range(2)]
    inner = pshape[2] * pshape[3]
    outer = pshape[1]
    calls = []

    def reader(sl, comp, s0, s1):
        calls.append((sl, comp, s0, s1))
        return pl[sl][comp, s0:s1]

    gpu.load_field_grid(reader, (nx, ny, nz), (0, 0, 0), (1, 1, 1), order=order,
                        two_slices=True, e0=0.25, b0=3.0, grid_dt=1.0,
                        stage_bytes=rows * 6 * inner * 4)
    bounds = list(range(0, outer, rows)) + [outer]
    want_calls = [(sl, c, bounds[k], bounds[k + 1]) for sl in (0, 1)
                  for k in range(len(bounds) - 1) for c in range(6)]
    assert calls == want_calls
    assert np.array_equal(gpu.field_nodes((nx, ny, nz), 0), expected_nodes(pl[0], order, 0.25, 3.0))
    assert np.array_equal(gpu.field_nodes((nx, ny, nz), 1), expected_nodes(pl[1], order, 0.25, 3.0))


def test_reader_failure_leaves_no_field(gpu):
    from test_particle_amd.gpu import TPError
    case("reader failure")
    gpu.set_field_test()

    def reader(sl, comp, s0, s1):
        return None if (comp == 4 and s0 > 0) else np.zeros((s1 - s0) * 6 * 5, np.float32)

    with pytest.raises(TPError) as ei:
        gpu.load_field_grid(reader, (5, 6, 8), (0, 0, 0), (1, 1, 1), stage_bytes=1)
    assert ei.value.code == abi.TP_E_ARG
    with pytest.raises(TPError) as ei:
        gpu.eval_field(np.zeros((1, 4)))
    assert ei.value.code == abi.TP_E_STATE
    with pytest.raises(TPError) as ei:
        gpu.field_nodes((5, 6, 8))
    assert ei.value.code == abi.TP_E_STATE


def test_bad_arguments(gpu):
    from test_particle_amd.gpu import TPError
    case("bad arguments")
    pl = np.zeros((6, 4, 4, 4), np.float32)
    for kw in [dict(order=2), dict(planes_prev=pl, grid_dt=0.0),
               dict(planes_prev=pl, grid_dt=float("nan")), dict(mode=5)]:
        with pytest.raises(TPError):
            gpu.set_field_grid_planes(pl, (0, 0, 0), (1, 1, 1), **kw)
    gpu.set_field_grid_planes(pl, (0, 0, 0), (1, 1, 1))
    with pytest.raises(TPError):
        gpu.field_nodes((4, 4, 5))                     # not the grid's dimensions
    with pytest.raises(TPError):
        gpu.field_nodes((4, 4, 4), 1)                  # one slice only


@pytest.mark.parametrize("two", [False, True])
@pytest.mark.parametrize("ct", [0, 2, 3])
def test_gda_files(gpu, tmp_path, two, ct):
    """tp_host_load_pic_gda streams the reference's .gda files (the golden
    slices) to the nodes tp_host_read_pic_gda + tp_gpu_set_field_grid(2)
    build, and those are the reference's loaders' arrays."""
    from test_particle_amd.gpu import read_pic_gda
    case(f"gda two={two} ct={ct}")
    g = golden("gda_slices")
    nx, ny, nz, nt = int(g["nx"]), int(g["ny"]), int(g["nz"]), int(g["nt"])
    for k in g.files:
        if k[:2] in ("ex", "ey", "ez", "bx", "by", "bz"):
            g[k].tofile(os.path.join(str(tmp_path), k + ".gda"))
    lo, hi = (0.0, 0.0, 0.0), (1.0, 2.0, 3.0)
    e0, b0 = float(g["e0"]), float(g["b0"])
    gpu.load_pic_gda(str(tmp_path), ct, (nx, ny, nz), lo, hi, e0, b0, nt=nt, two_slices=two,
                     grid_dt=0.5)
    ref = g[f"mt{int(two)}_ct{ct}"].reshape(-1, nz, ny, nx, 3)
    # the reference's order: E then B of the next (or only) slice, then the previous
    got = [gpu.field_nodes((nx, ny, nz), s) for s in range(1 + int(two))]
    want = [ref[0], ref[1]] + ([ref[2], ref[3]] if two else [])
    for s, nodes in enumerate(got):
        assert np.array_equal(nodes[..., :3], want[2 * s])
        assert np.array_equal(nodes[..., 3:6], want[2 * s + 1])
        assert not nodes[..., 6:].any()
    arrs = read_pic_gda(str(tmp_path), ct, nx, ny, nz, e0, b0, nt=nt, two_slices=two)
    assert all(np.array_equal(a, w) for a, w in zip(arrs, want))


def test_gda_missing_or_short_file(gpu, tmp_path):
    from test_particle_amd.gpu import TPError
    case("gda errors")
    g = golden("gda_slices")
    nx, ny, nz = int(g["nx"]), int(g["ny"]), int(g["nz"])
    for k in g.files:
        if k.endswith("_0") and k[:2] in ("ex", "ey", "ez", "bx", "by", "bz"):
            g[k].tofile(os.path.join(str(tmp_path), k + ".gda"))
    with pytest.raises(TPError):
        gpu.load_pic_gda(str(tmp_path), 1, (nx, ny, nz), (0, 0, 0), (1, 1, 1), 1.0, 1.0)
    g["bz_0"][:-1].tofile(os.path.join(str(tmp_path), "bz_0.gda"))
    with pytest.raises(TPError):
        gpu.load_pic_gda(str(tmp_path), 0, (nx, ny, nz), (0, 0, 0), (1, 1, 1), 1.0, 1.0)


def test_h5_file_equals_reference_loaders(gpu, tmp_path):
    """tp_h5_load_pic streams the HDF5 file (hyperslabs of its larger
    datasets) to the arrays the reference's read_pic_*fields_h5 built."""
    from test_particle_amd import h5
    case("h5 golden")
    g = golden("pic_h5")
    nx, ny, nz = int(g["nx"]), int(g["ny"]), int(g["nz"])
    fname = str(tmp_path / "fields_3.h5")
    h5.write_pic_planes(fname, str(g["gname"]), g["planes"])
    lo, hi = (0.0, 0.0, 0.0), (6.0, 4.0, 5.0)
    h5.load_pic(gpu, fname, str(g["gname"]), (nx, ny, nz), lo, hi, float(g["e0"]),
                float(g["b0"]))
    nodes = gpu.field_nodes((nx, ny, nz))
    assert np.array_equal(nodes[..., :3], g["e"])
    assert np.array_equal(nodes[..., 3:6], g["b"])
    pts = rand_points(np.random.default_rng(3), lo, hi, 300)
    f_h5 = gpu.eval_field(pts)
    gpu.set_field_grid(g["e"], g["b"], lo, hi)
    assert np.array_equal(gpu.eval_field(pts), f_h5)
    from test_particle_amd.gpu import TPError
    with pytest.raises(TPError):
        h5.load_pic(gpu, fname, "Timestep_9", (nx, ny, nz), lo, hi, 1.0, 1.0)
    with pytest.raises(TPError):
        h5.load_pic(gpu, fname, str(g["gname"]), (nx, 7, nz), lo, hi, 1.0, 1.0)


@pytest.mark.parametrize("order", [abi.TP_PLANES_X_FASTEST, abi.TP_PLANES_X_SLOWEST])
def test_large_grid(gpu, order):
    """256 x 192 x 160 nodes (several 32 MiB slabs): the nodes equal the
    closed form."""
    nx, ny, nz = 256, 192, 160
    case(f"large grid order={order}")
    rng = np.random.default_rng(8)
    pshape = (6, nz, ny, nx) if order == abi.TP_PLANES_X_FASTEST else (6, nx, ny, nz)
    planes = rng.standard_normal(pshape, dtype=np.float32)
    gpu.set_field_grid_planes(planes, (0, 0, 0), (1, 1, 1), order=order, e0=0.3, b0=1.7)
    got = gpu.field_nodes((nx, ny, nz))
    assert np.array_equal(got, expected_nodes(planes, order, 0.3, 1.7))
