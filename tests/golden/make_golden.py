"""Write the golden vectors of tests/golden/ from the reference tracer.

Builds oracle/_ref (the reference's own hot-path sources compiled from
/root/reference/src by oracle/Makefile, linked to oracle/ref_harness.c) and
runs it on seeded inputs.  Every fixture is an .npz of inputs and the
reference's outputs; nothing of the reference's source is stored.

    python tests/golden/make_golden.py      # in the build container only
    python tests/golden/make_golden.py h5   # only the PIC HDF5 input fixture
    python tests/golden/make_golden.py mhd  # only the system-3 and double-grid fixtures
"""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, "..", ".."))
sys.path.insert(0, ROOT)

from test_particle_amd.particles import init_particles  # noqa: E402
from test_particle_amd.poincare import write_config  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref")
TWO_PI = 2.0 * np.pi


def build_ref():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"],
                   check=True)


def write_spec(path, **kw):
    with open(path, "w") as f:
        for k, v in kw.items():
            if isinstance(v, (float, np.floating)):
                f.write(f"{k} {float(v)!r}\n")
            else:
                f.write(f"{k} {v}\n")


def run(binary, *args, cwd):
    subprocess.run([os.path.join(REF, binary), *map(str, args)], cwd=cwd,
                   check=True, stdout=subprocess.DEVNULL)


def box(lo, hi):
    return dict(xmin=float(lo[0]), ymin=float(lo[1]), zmin=float(lo[2]),
                xmax=float(hi[0]), ymax=float(hi[1]), zmax=float(hi[2]))


def fields_case(name, binary, spec, pts, extra=None, tmp=None):
    sp = os.path.join(tmp, "spec.txt")
    write_spec(sp, **spec)
    pts.astype(np.float64).tofile(os.path.join(tmp, "pts.bin"))
    run(binary, "fields", sp, "pts.bin", "out.bin", cwd=tmp)
    out = np.fromfile(os.path.join(tmp, "out.bin")).reshape(-1, 6)
    np.savez(os.path.join(HERE, name + ".npz"), pts=pts, out=out,
             **(extra or {}))


def wirz_case(name, binary, spec, ptl, ks, extra=None, tmp=None):
    sp = os.path.join(tmp, "spec.txt")
    write_spec(sp, **spec)
    ptl.tofile(os.path.join(tmp, "ptl.bin"))
    res = {}
    for k in ks:
        run(binary, "wirz", sp, "ptl.bin", k, "out.bin", cwd=tmp)
        res[f"out_{k}"] = np.fromfile(os.path.join(tmp, "out.bin")).reshape(-1, 11)
    np.savez(os.path.join(HERE, name + ".npz"), ptl=ptl, ks=np.array(ks),
             **res, **(extra or {}))


def track_case(name, binary, spec, ptl, extra=None, tmp=None):
    sp = os.path.join(tmp, "spec.txt")
    write_spec(sp, **spec)
    ptl.tofile(os.path.join(tmp, "ptl.bin"))
    od = os.path.join(tmp, "o")
    os.makedirs(od, exist_ok=True)
    run(binary, "track", sp, "ptl.bin", od, cwd=tmp)
    rd = lambda f, dt=np.float64: np.fromfile(os.path.join(od, f), dtype=dt)
    nb, nt = spec["nbins"], spec["nt_out"]
    np.savez(os.path.join(HERE, name + ".npz"), ptl_in=ptl,
             ptl_out=rd("ptl.bin").reshape(-1, 11),
             nsteps=rd("nsteps.bin", np.int32),
             espect=rd("espect.bin").reshape(nt, nb),
             espect_escape=rd("espect_escape.bin").reshape(nt, nb),
             dcoeff=rd("dcoeff.bin").reshape(8, -1),
             remain=rd("remain.bin", np.int32),
             spec=np.array(json.dumps({k: float(v) if isinstance(v, np.floating) else v
                                       for k, v in spec.items()})),
             **(extra or {}))


def traj_case(name, binary, spec, ptl, extra=None, tmp=None):
    """The fixed-step loop with trajectory diagnostics: every particle has
    estep/nsteps_output + 1 slots, slot 0 its initial state."""
    sp = os.path.join(tmp, "spec.txt")
    write_spec(sp, **spec)
    ptl.tofile(os.path.join(tmp, "ptl.bin"))
    od = os.path.join(tmp, "t")
    os.makedirs(od, exist_ok=True)
    run(binary, "traj", sp, "ptl.bin", od, cwd=tmp)
    rd = lambda f, dt=np.float64: np.fromfile(os.path.join(od, f), dtype=dt)
    nb, nt = spec["nbins"], spec["nt_out"]
    np.savez(os.path.join(HERE, name + ".npz"), ptl_in=ptl,
             ptl_out=rd("ptl.bin").reshape(-1, 11),
             traj=rd("traj.bin").reshape(ptl.shape[0], -1, 11),
             points=rd("points.bin", np.int32),
             espect=rd("espect.bin").reshape(nt, nb),
             spec=np.array(json.dumps({k: float(v) if isinstance(v, np.floating) else v
                                       for k, v in spec.items()})),
             **(extra or {}))


PMAP_KEYS = ("face", "step", "length_max", "ncross", "single", "x0", "y0", "z0", "min_h",
             "max_h", "min_v", "max_v", "nh", "nv")


def pmap_case(name, spec, cfg, extra, tmp):
    """Poincare map of the reference's next_location and config parser
    (ref_harness mode pmap)."""
    sp = os.path.join(tmp, "spec.txt")
    write_spec(sp, **spec)
    cp = os.path.join(tmp, "pmap.cfg")
    write_config(cp, cfg)
    run("ref_harness", "pmap", sp, cp, "pm.bin", cwd=tmp)
    np_ = 1 if cfg["single"] else cfg["nh"] * cfg["nv"]
    out = np.fromfile(os.path.join(tmp, "pm.bin")).reshape(3, np_, cfg["ncross"] + 1)
    np.savez(os.path.join(HERE, name + ".npz"), xpos=out[0], ypos=out[1], zpos=out[2],
             cfg=np.array(json.dumps({k: cfg[k] for k in PMAP_KEYS})),
             **extra)


def h5_case(tmp):
    """PIC HDF5 field input: a file of the reference's layout (datasets ex ey
    ez cbx cby cbz, x slowest, larger than the block read) written with
    libtph5.so, loaded by the reference's read_pic_efields_h5 and
    read_pic_bfields_h5."""
    import ctypes
    from test_particle_amd import build
    build.build_h5()
    L = ctypes.CDLL(build.OUT_H5)
    rng = np.random.default_rng(91)
    nx, ny, nz = 7, 5, 6
    dims = np.array([9, 6, 8], dtype=np.int64)        # (x, y, z) extents in the file
    planes = rng.normal(0, 1, (6, *dims)).astype(np.float32)
    fname = os.path.join(tmp, "fields_3.h5")
    rc = L.tp_h5_write_pic_planes(fname.encode(), b"Timestep_3",
                                  dims.ctypes.data_as(ctypes.c_void_p),
                                  planes.ctypes.data_as(ctypes.c_void_p))
    assert rc == 0
    sp = os.path.join(tmp, "spec.txt")
    write_spec(sp, system=4, nx=nx, ny=ny, nz=nz, tmax=1.0, dt=0.01, gda_b0=2.5, gda_e0=0.3,
               **box((0.0, 0.0, 0.0), (1.0, 1.0, 1.0)))
    run("ref_harness", "h5", sp, fname, "Timestep_3", "hout.bin", cwd=tmp)
    out = np.fromfile(os.path.join(tmp, "hout.bin"), dtype=np.float32).reshape(2, nz, ny, nx, 3)
    np.savez(os.path.join(HERE, "pic_h5.npz"), nx=nx, ny=ny, nz=nz, e0=0.3, b0=2.5,
             planes=planes, gname="Timestep_3", e=out[0], b=out[1])


def mhd_cases(tmp):
    """System 3 (MHD: float velocity and magnetic grids, E = B x v,
    src/emfields.c:107-120) through get_emf, tracking_wirz and the
    fixed-step loop; and the reference's double gathers
    (src/electric_field.c:370-406, src/magnetic_field.c:555-591) at points."""
    os.makedirs(os.path.join(tmp, "data"), exist_ok=True)
    rng = np.random.default_rng(313)
    glo, ghi = (0.0, -1.0, 2.0), (5.0, 3.0, 5.0)
    nz, ny, nx = 4, 5, 6
    vg = rng.normal(0, 0.2, (nz, ny, nx, 3)).astype(np.float32)
    bg = rng.normal(0, 1, (nz, ny, nx, 3)).astype(np.float32)
    bg[..., 2] += 2.0
    np.concatenate([vg.ravel(), bg.ravel()]).tofile(os.path.join(tmp, "mgrid.bin"))
    mspec = dict(system=3, nx=nx, ny=ny, nz=nz, tmax=1.0, dt=0.01,
                 grid=os.path.join(tmp, "mgrid.bin"), **box(glo, ghi))
    mp = np.column_stack([rng.uniform(-0.3, 5.3, 400), rng.uniform(-1.2, 3.2, 400),
                          rng.uniform(1.8, 5.2, 400), np.zeros(400)])
    fields_case("fields_mhd", "ref_harness", mspec, mp,
                extra=dict(vgrid=vg, bgrid=bg, lo=glo, hi=ghi), tmp=tmp)
    mptl = init_particles(32, glo, ghi, 0.3, seed=314)
    wirz_case("wirz_mhd", "ref_harness", dict(mspec, bc=0), mptl, [1, 50, 200],
              extra=dict(dt=0.01, vgrid=vg, bgrid=bg, lo=glo, hi=ghi), tmp=tmp)
    track_case("track_mhd", "ref_harness",
               dict(mspec, tmax=300 * 0.01 + 1e-9, bc=0, nbins=50, nt_out=3,
                    emin=1e-3, emax=1.0), mptl,
               extra=dict(vgrid=vg, bgrid=bg, lo=glo, hi=ghi, dt=0.01), tmp=tmp)
    # double grids: the gathers alone (the reference's get_emf never calls them)
    eg = rng.normal(0, 1, (nz, ny, nx, 3))
    bd = rng.normal(0, 1, (nz, ny, nx, 3))
    np.concatenate([eg.ravel(), bd.ravel()]).tofile(os.path.join(tmp, "dgrid.bin"))
    dspec = dict(system=0, nx=nx, ny=ny, nz=nz, tmax=1.0, dt=0.01,
                 dgrid=os.path.join(tmp, "dgrid.bin"), **box(glo, ghi))
    sp = os.path.join(tmp, "spec.txt")
    write_spec(sp, **dspec)
    mp.tofile(os.path.join(tmp, "pts.bin"))
    run("ref_harness", "dfields", sp, "pts.bin", "out.bin", cwd=tmp)
    out = np.fromfile(os.path.join(tmp, "out.bin")).reshape(-1, 6)
    np.savez(os.path.join(HERE, "fields_gridd.npz"), pts=mp, out=out, egrid=eg, bgrid=bd,
             lo=glo, hi=ghi)


def main():
    build_ref()
    if sys.argv[1:] == ["h5"]:
        with tempfile.TemporaryDirectory() as tmp:
            h5_case(tmp)
        return
    if sys.argv[1:] == ["mhd"]:
        with tempfile.TemporaryDirectory() as tmp:
            mhd_cases(tmp)
        return
    rng = np.random.default_rng(7)
    lo2pi, hi2pi = (0.0, 0.0, 0.0), (TWO_PI, TWO_PI, TWO_PI)
    with tempfile.TemporaryDirectory() as tmp:
        os.makedirs(os.path.join(tmp, "data"))
        # --- ABC field, reference and unit coefficients ------------------
        pts = np.concatenate([rng.uniform(-1.0, 8.0, (200, 4)),
                              np.zeros((1, 4))])
        base = dict(system=2, nx=5, ny=5, nz=5, tmax=1.0, dt=0.01, **box(lo2pi, hi2pi))
        fields_case("fields_abc_ref", "ref_harness", base, pts, tmp=tmp)
        fields_case("fields_abc1", "ref_harness_abc1",
                    dict(base, ff_b0=1.3, ff_lambda=0.7), pts,
                    extra=dict(b0=1.3, lam=0.7), tmp=tmp)
        # --- ABC + E = B x v from a random double velocity grid -----------
        vlo, vhi = (-0.5, 0.25, 1.0), (3.5, 2.25, 4.0)
        vg = rng.normal(0.0, 0.1, (4, 6, 7, 3))      # (nz, ny, nx, 3)
        vg.tofile(os.path.join(tmp, "vgrid.bin"))
        vpts = np.column_stack([rng.uniform(-0.8, 3.8, 300), rng.uniform(0.0, 2.5, 300),
                                rng.uniform(0.8, 4.2, 300), np.zeros(300)])
        fields_case("fields_abc_vgrid", "ref_harness",
                    dict(system=2, nx=7, ny=6, nz=4, tmax=1.0, dt=0.01,
                         vgrid=os.path.join(tmp, "vgrid.bin"), **box(vlo, vhi)),
                    vpts, extra=dict(vgrid=vg, lo=vlo, hi=vhi), tmp=tmp)
        # --- PIC float grid (system 4), x/z weight exchange and wraps ------
        glo, ghi = (0.0, -1.0, 2.0), (5.0, 3.0, 5.0)
        nz, ny, nx = 4, 5, 6
        eg = rng.normal(0, 1, (nz, ny, nx, 3)).astype(np.float32)
        bg = rng.normal(0, 1, (nz, ny, nx, 3)).astype(np.float32)
        np.concatenate([eg.ravel(), bg.ravel()]).tofile(os.path.join(tmp, "grid.bin"))
        gp = np.column_stack([rng.uniform(-0.3, 5.3, 400), rng.uniform(-1.2, 3.2, 400),
                              rng.uniform(1.8, 5.2, 400), np.zeros(400)])
        nodes = np.array([[1.0, 0.0, 3.0, 0.0], [0.0, -1.0, 2.0, 0.0],
                          [2.5, 1.5, 3.5, 0.0]])
        gp = np.concatenate([gp, nodes])
        gspec = dict(system=4, nx=nx, ny=ny, nz=nz, tmax=1.0, dt=0.01,
                     grid=os.path.join(tmp, "grid.bin"), **box(glo, ghi))
        fields_case("fields_grid", "ref_harness", gspec, gp,
                    extra=dict(egrid=eg, bgrid=bg, lo=glo, hi=ghi), tmp=tmp)
        # --- WLCS (static wire and Bessel wire) ---------------------------
        rows = np.array([
            [1.0, 0.5, 0.5, 0.5, 0.0, 0.0, 0.0, 0.0, 1.0, 0.5, 0.5, 0.5, 0.0, 0.0],
            [0.7, 0.2, 0.9, 0.4, 0.3, 1.1, 10.0, 0.6, 0.3, 0.8, 0.1, 0.6, 0.5, 2.0],
            [-0.4, 0.8, 0.1, 0.7, 2.0, 4.0, 5.0, 1.2, 0.7, 0.3, 0.6, 0.2, 1.3, 0.4]])
        rows.tofile(os.path.join(tmp, "wlcs.bin"))
        wp = np.column_stack([rng.uniform(-0.5, 1.5, (300, 3)), rng.uniform(0, 50, 300)])
        for ib in (0, 1):
            fields_case(f"fields_wlcs_b{ib}", "ref_harness",
                        dict(system=1, nwlcs=3, ibessel=ib, wlcs_omega=0.001,
                             wlcs_normi=0.125, wlcs=os.path.join(tmp, "wlcs.bin"),
                             tmax=1.0, dt=0.01, **box((-1, -1, -1), (2, 2, 2))),
                        wp, extra=dict(rows=rows, ibessel=ib, omega=0.001,
                                       norm_i=0.125), tmp=tmp)
        # --- Wirz trajectories ---------------------------------------------
        ptl = init_particles(64, lo2pi, hi2pi, 0.7, seed=11)
        wirz_case("wirz_test", "ref_harness",
                  dict(system=0, tmax=1.0, dt=0.05, bc=0, **box(lo2pi, hi2pi)),
                  ptl, [1, 1000], extra=dict(dt=0.05), tmp=tmp)
        dt_c1 = 0.01 * TWO_PI * (1.0 / np.sqrt(1.0 - 0.49)) / np.sqrt(3.0)
        wirz_case("wirz_abc1", "ref_harness_abc1",
                  dict(system=2, nx=5, ny=5, nz=5, tmax=1.0, dt=dt_c1, bc=0,
                       **box(lo2pi, hi2pi)),
                  ptl, [1, 10, 100, 1000], extra=dict(dt=dt_c1), tmp=tmp)
        gptl = init_particles(32, glo, ghi, 0.3, seed=12)
        wirz_case("wirz_grid", "ref_harness", dict(gspec, bc=0), gptl, [1, 50, 200],
                  extra=dict(dt=0.01, egrid=eg, bgrid=bg, lo=glo, hi=ghi), tmp=tmp)
        # --- calc_diff_coeff ------------------------------------------------
        p0 = init_particles(64, lo2pi, hi2pi, 0.7, seed=13)
        p1 = p0.copy()
        p1[:, :3] += rng.normal(0, 0.5, (64, 3))
        p1[:, 4:7] += rng.normal(0, 0.1, (64, 3))
        p1[:, 7] = 3.7
        p1[:, 8:11] = rng.choice([-TWO_PI, 0.0, TWO_PI], (64, 3))
        sp = os.path.join(tmp, "spec.txt")
        write_spec(sp, system=2, nx=5, ny=5, nz=5, tmax=1.0, dt=0.01, **box(lo2pi, hi2pi))
        p0.tofile(os.path.join(tmp, "p0.bin"))
        p1.tofile(os.path.join(tmp, "p1.bin"))
        run("ref_harness_abc1", "diff", sp, "p0.bin", "p1.bin", 0.05, "d.bin", cwd=tmp)
        np.savez(os.path.join(HERE, "diff_abc1.npz"), p0=p0, p1=p1, dt=0.05,
                 out=np.fromfile(os.path.join(tmp, "d.bin")).reshape(-1, 8))
        # --- full fixed-step loop ------------------------------------------
        tptl = init_particles(256, lo2pi, hi2pi, 0.7, seed=14)
        track_case("track_abc1", "ref_harness_abc1",
                   dict(system=2, nx=5, ny=5, nz=5, tmax=1000 * dt_c1 + 1e-9,
                        dt=dt_c1, bc=0, nbins=102, nt_out=10, emin=1e-3, emax=10.0,
                        **box(lo2pi, hi2pi)), tptl, tmp=tmp)
        # open boundaries in the uniform test field: particles stream out
        olo, ohi = (0.0, 0.0, 0.0), (4.0, 4.0, 4.0)
        optl = init_particles(256, olo, ohi, 0.5, seed=15)
        optl[-8:, 2] = 4.5         # a few start outside the box
        optl[0, :3] = 2.0          # particle 0 stays inside until its first
        optl[0, 6] = 0.0           # D sample (the reference reads stale terms)
        track_case("track_open", "ref_harness",
                   dict(system=0, tmax=600 * 0.05 + 1e-9, dt=0.05, bc=1,
                        nbins=30, nt_out=6, emin=0.01, emax=1.0,
                        **box(olo, ohi)), optl, tmp=tmp)
        # PIC grid loop
        track_case("track_grid", "ref_harness",
                   dict(gspec, tmax=300 * 0.01 + 1e-9, bc=0, nbins=50, nt_out=3,
                        emin=1e-3, emax=1.0), gptl,
                   extra=dict(egrid=eg, bgrid=bg, lo=glo, hi=ghi, dt=0.01), tmp=tmp)
        # --- trajectory diagnostics (traj_diagnose = 1) --------------------
        traj_case("traj_grid", "ref_harness",
                  dict(gspec, tmax=300 * 0.01 + 1e-9, bc=0, nbins=50, nt_out=3,
                       emin=1e-3, emax=1.0, nsteps_output=7), gptl,
                  extra=dict(egrid=eg, bgrid=bg, lo=glo, hi=ghi, dt=0.01), tmp=tmp)
        traj_case("traj_open", "ref_harness",
                  dict(system=0, tmax=600 * 0.05 + 1e-9, dt=0.05, bc=1,
                       nbins=30, nt_out=6, emin=0.01, emax=1.0, nsteps_output=25,
                       **box(olo, ohi)), optl[:64].copy(), tmp=tmp)
        # --- PIC grid with two time slices (multi_tframe = 1) --------------
        rng2 = np.random.default_rng(99)
        slices = [rng2.normal(0, 1, (nz, ny, nx, 3)).astype(np.float32) for _ in range(4)]
        slices[1][..., 2] += 1.5                 # B of slice a
        slices[3][..., 2] += 1.5                 # B of slice b
        np.concatenate([a.ravel() for a in slices]).tofile(os.path.join(tmp, "grid2.bin"))
        g2spec = dict(gspec, mt=1, grid_dt=0.37, grid=os.path.join(tmp, "grid2.bin"))
        g2p = np.column_stack([rng2.uniform(-0.3, 5.3, 400), rng2.uniform(-1.2, 3.2, 400),
                               rng2.uniform(1.8, 5.2, 400), rng2.uniform(0.0, 3.0, 400)])
        fields_case("fields_grid2", "ref_harness", g2spec, g2p,
                    extra=dict(ea=slices[0], ba=slices[1], eb=slices[2], bb=slices[3],
                               lo=glo, hi=ghi, grid_dt=0.37), tmp=tmp)
        g2ptl = init_particles(32, glo, ghi, 0.3, seed=17)
        track_case("track_grid2", "ref_harness",
                   dict(g2spec, tmax=300 * 0.01 + 1e-9, bc=0, nbins=50, nt_out=3,
                        emin=1e-3, emax=1.0), g2ptl,
                   extra=dict(ea=slices[0], ba=slices[1], eb=slices[2], bb=slices[3],
                              lo=glo, hi=ghi, grid_dt=0.37, dt=0.01), tmp=tmp)
        # --- PIC field input from .gda files (one and two time slices) -----
        rng3 = np.random.default_rng(77)
        gdir = os.path.join(tmp, "gda") + "/"
        os.makedirs(gdir, exist_ok=True)
        files = {}
        for ct in range(4):
            for comp in ("ex", "ey", "ez", "bx", "by", "bz"):
                a = rng3.normal(0, 1, nz * ny * nx).astype(np.float32)
                a.tofile(f"{gdir}{comp}_{ct}.gda")
                files[f"{comp}_{ct}"] = a
        out = {}
        for mt in (0, 1):
            for ct in (0, 2, 3):
                sp = os.path.join(tmp, "spec.txt")
                write_spec(sp, **dict(gspec, mt=mt, grid_nt=4, grid_dt=0.5, gda_b0=2.5,
                                      gda_e0=0.3))
                run("ref_harness", "gda", sp, gdir, ct, "gout.bin", cwd=tmp)
                out[f"mt{mt}_ct{ct}"] = np.fromfile(os.path.join(tmp, "gout.bin"),
                                                    dtype=np.float32)
        np.savez(os.path.join(HERE, "gda_slices.npz"), nx=nx, ny=ny, nz=nz, nt=4, b0=2.5,
                 e0=0.3, **files, **out)
        h5_case(tmp)
        # --- Poincare map (run_type 2) --------------------------------------
        rng4 = np.random.default_rng(31)
        plo, phi = (0.0, -1.0, 2.0), (6.0, 4.0, 5.0)
        pnz, pny, pnx = 8, 10, 12
        # a guide field along y with perturbations: lines cross y planes
        pb = (0.35 * rng4.normal(0, 1, (pnz, pny, pnx, 3))).astype(np.float32)
        pb[..., 1] += 1.0
        pe = np.zeros_like(pb)
        np.concatenate([pe.ravel(), pb.ravel()]).tofile(os.path.join(tmp, "pgrid.bin"))
        pspec = dict(system=4, nx=pnx, ny=pny, nz=pnz, tmax=1.0, dt=0.01,
                     grid=os.path.join(tmp, "pgrid.bin"), **box(plo, phi))
        pext = dict(bgrid=pb, lo=plo, hi=phi)
        pmap_case("pmap_grid", pspec,
                  dict(face=1, step=0.05, length_max=27.0, ncross=12, single=0, x0=0.0,
                       y0=1.3, z0=0.0, min_h=0.2, max_h=5.7, min_v=2.1, max_v=4.9, nh=5, nv=4),
                  pext, tmp)
        # a single point on the box face z = zmin, crossing z planes
        pmap_case("pmap_single", pspec,
                  dict(face=2, step=0.04, length_max=80.0, ncross=25, single=1, x0=3.1,
                       y0=0.4, z0=2.0, min_h=0.0, max_h=0.0, min_v=0.0, max_v=0.0, nh=1,
                       nv=1), pext, tmp)
        # one row of x planes (nv = 1), field of the two-slice setup at t = 0
        pb2 = [(0.3 * rng4.normal(0, 1, (pnz, pny, pnx, 3))).astype(np.float32)
               for _ in range(2)]
        pb2[0][..., 2] += 1.0                    # slice a (next): guide along z
        pb2[1][..., 0] += 1.0                    # slice b (previous): guide along x
        np.concatenate([pe.ravel(), pb2[0].ravel(), pe.ravel(),
                        pb2[1].ravel()]).tofile(os.path.join(tmp, "pgrid2.bin"))
        pmap_case("pmap_grid2", dict(pspec, mt=1, grid_dt=0.5,
                                     grid=os.path.join(tmp, "pgrid2.bin")),
                  dict(face=0, step=0.05, length_max=24.0, ncross=10, single=0, x0=2.2,
                       y0=0.0, z0=0.0, min_h=-0.8, max_h=3.6, min_v=3.3, max_v=3.3, nh=7,
                       nv=1),
                  dict(b_next=pb2[0], b_prev=pb2[1], lo=plo, hi=phi, grid_dt=0.5), tmp)
        # --- the reference's own output files (format) ---------------------
        sp = os.path.join(tmp, "spec.txt")
        write_spec(sp, system=2, nx=5, ny=5, nz=5, tmax=300 * dt_c1 + 1e-9,
                   dt=dt_c1, bc=0, nbins=102, nt_out=3, emin=1e-3, emax=10.0,
                   **box(lo2pi, hi2pi))
        hptl = init_particles(64, lo2pi, hi2pi, 0.7, seed=16)
        hptl.tofile(os.path.join(tmp, "h.bin"))
        run("ref_harness_abc1", "hybrid", sp, "h.bin", cwd=tmp)
        with open(os.path.join(tmp, "data", "espectrum.dat")) as f:
            espectrum_dat = f.read()
        with open(os.path.join(tmp, "data", "diff_coeffs.dat")) as f:
            diff_dat = f.read()
        np.savez(os.path.join(HERE, "hybrid_files_abc1.npz"), ptl=hptl,
                 dt=dt_c1, espectrum_dat=np.array(espectrum_dat),
                 diff_coeffs_dat=np.array(diff_dat))
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    main()
