"""The allocation-free run (tp_gpu_run_fixed_noalloc) against tp_gpu_run_fixed.

tp_gpu_run_fixed_noalloc runs the same kernels as tp_gpu_run_fixed, with its
scratch (D partial-sum rows, open-boundary index lists, the non-finite flag)
in the library's static device array and its outputs in device buffers the
caller owns.  Two calls in a row on one context, the second over fewer
particles, must give what tp_gpu_run_fixed gives on the same particles:
particles, step counts, spectra, remaining counts bit for bit, and the D
sums bit for bit (same blocks, same fixed-order reduction).  The output
buffers start with garbage, which the call must clear."""
import numpy as np
import pytest

from test_particle_amd import abi
from test_particle_amd.particles import init_particles

pytestmark = pytest.mark.gpu

TWO_PI = 2.0 * np.pi
BOX = ((0.0, 0.0, 0.0), (TWO_PI, TWO_PI, TWO_PI))
DT = 0.01 * TWO_PI * (1.0 / np.sqrt(1.0 - 0.49)) / np.sqrt(3.0)


@pytest.fixture(scope="module")
def gpu():
    from test_particle_amd.gpu import Tracker
    t = Tracker(0)
    yield t
    t.close()


def _reference_run(t, prm, ptl):
    t.load(ptl)
    t.run(prm)
    p, ns = t.store()
    return dict(ptl=p, nsteps=ns, **t.diagnostics(prm))


def _noalloc_run(t, prm, ptl):
    from hipmem import DeviceArray
    nd = abi.n_dcoeff_rows(prm)
    nbt = prm.nt_out * prm.nbins
    # the caller's device buffers, holding garbage the call has to clear
    esp = DeviceArray(np.full(nbt, 7, np.uint32))
    ese = DeviceArray(np.full(nbt, 9, np.uint32))
    dco = DeviceArray(np.full(8 * nd, 3.5, np.float64))
    rem = DeviceArray(np.full(nd, -11, np.int32))
    t.load(ptl)
    t.run_noalloc(prm, esp.address, ese.address, dco.address, rem.address)
    p, ns = t.store()
    out = dict(ptl=p, nsteps=ns,
               espect=esp.download().astype(np.float64).reshape(prm.nt_out, prm.nbins),
               espect_escape=ese.download().astype(np.float64).reshape(prm.nt_out, prm.nbins),
               dcoeff=dco.download().reshape(8, nd), remain=rem.download())
    for b in (esp, ese, dco, rem):
        b.free()
    return out


def _same(a, b, what):
    for k in ("ptl", "nsteps", "espect", "espect_escape", "dcoeff", "remain"):
        x, y = np.asarray(a[k]), np.asarray(b[k])
        assert x.shape == y.shape, f"{what} {k}: shape {x.shape} vs {y.shape}"
        assert np.array_equal(x, y), f"{what} {k}: {np.count_nonzero(x != y)} entries differ"


def _case(n, estep, nt_out, bc, pusher, seed):
    ptl = init_particles(n, BOX[0], BOX[1], 0.7, seed=seed)
    prm = abi.make_params(dt=DT, estep=estep, lo=BOX[0], hi=BOX[1], pusher=pusher,
                          bc_flag=bc, nbins=102, nt_out=nt_out, emin=1e-3, emax=10.0,
                          flags=abi.TP_FLAG_CHECK_FINITE)
    return prm, ptl


def test_noalloc_matches_run_fixed_two_calls(gpu):
    gpu.set_field_abc(1.0, 1.0, 1.0)
    # open boundaries: after the first escapes the launches run over the
    # index lists held in the scratch array; then a smaller periodic run
    cases = [_case(5000, 500, 5, abi.TP_BC_OPEN, abi.TP_PUSH_BORIS, 11),
             _case(700, 300, 3, abi.TP_BC_PERIODIC, abi.TP_PUSH_WIRZ, 12)]
    for k, (prm, ptl) in enumerate(cases):
        print(f"[noalloc] call {k}: n={ptl.shape[0]} estep={prm.estep} bc={prm.bc_flag}",
              flush=True)
        ref = _reference_run(gpu, prm, ptl)
        got = _noalloc_run(gpu, prm, ptl)
        _same(got, ref, f"call {k}")
        if prm.bc_flag == abi.TP_BC_OPEN:
            assert ref["remain"][-1] < ptl.shape[0]       # some particles escaped
            assert ref["espect_escape"].sum() > 0


def test_noalloc_refuses_missing_outputs(gpu):
    from test_particle_amd.gpu import TPError
    gpu.set_field_abc(1.0, 1.0, 1.0)
    prm, ptl = _case(64, 10, 1, abi.TP_BC_PERIODIC, abi.TP_PUSH_BORIS, 13)
    gpu.load(ptl)
    with pytest.raises(TPError) as e:
        gpu.run_noalloc(prm, 0, 0, 0, 0)
    assert e.value.code == abi.TP_E_ARG
