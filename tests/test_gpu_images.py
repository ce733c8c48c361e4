"""GPU parity of the old imager and the optical-depth images
(mgx_image_ray_device, mgx_image_optical_depth_ray_device,
mgx_image_optical_depth_device) against the oracle's restatement of
Solver::image_feautrier_order_2(model, rr), image_optical_depth(model, rr) and
image_optical_depth_new_imager on the same models: FP64 within 1e-10 relative
(floor 1e-16 of the largest value), ImX and ImY bit-exact; mixed precision
within 1e-5 of the FP64 oracle."""
import numpy as np
import pytest

from magritte_amd import _abi

pytestmark = pytest.mark.gpu

RTOL = 1e-10


def _dev(model):
    from magritte_amd.solver import DeviceModel

    return DeviceModel(model)


def _image_case(m, nfreq=31):
    lo = float(m.line[0]) * (1 - 2e-4)
    hi = float(m.line[0]) * (1 + 2e-4)
    return m.image_frequencies(lo, hi, nfreq)


def _err(got, ref):
    scale = np.max(np.abs(ref))
    return float(np.max(np.abs(got - ref) / np.maximum(np.abs(ref), 1e-6 * scale)))


# (model, ray, approximation, image grid): the 1D benchmark along its (0, 1, 0) ray;
# the 12^3 wind along a ray below hnrays and one above (reversed shifts)
CASES = [("vgrad_model", -1, _abi.APPROX_CLOSE_LINES, True),
         ("imager_model", -1, _abi.APPROX_CLOSE_LINES, True),
         ("small_wind", 0, _abi.APPROX_CLOSE_LINES, False),
         ("small_wind", 8, _abi.APPROX_NONE, True),
         ("small_wind", 3, _abi.APPROX_CLOSE_LINES, True)]


@pytest.mark.parametrize("which,rr,approx,grid", CASES)
@pytest.mark.parametrize("optical_depth", [False, True])
def test_image_ray_fp64_parity(which, rr, approx, grid, optical_depth, request):
    from magritte_amd import solver
    from oracle import oracle

    m = request.getfixturevalue(which)
    rr = m.hnrays - 1 if rr < 0 else rr
    op = m.options(approx=approx)
    fq = _image_case(m) if grid else None
    print(f"[image_ray] model={which} rr={rr} approx={approx} grid={grid} "
          f"optical_depth={optical_depth}", flush=True)
    out, X, Y = solver.image_ray(_dev(m), op, rr, fq, optical_depth=optical_depth)
    out, X, Y = out.cpu().numpy(), X.cpu().numpy(), Y.cpu().numpy()
    ref, Xr, Yr = oracle.image_ray(m, op, rr, fq, optical_depth=optical_depth)
    np.testing.assert_array_equal(X.view(np.uint64), Xr.view(np.uint64))
    np.testing.assert_array_equal(Y.view(np.uint64), Yr.view(np.uint64))
    assert np.isfinite(out).all()
    err = _err(out, ref)
    print(f"  max rel err {err:.3e}", flush=True)
    assert err < RTOL


@pytest.mark.parametrize("which,raydir,nx,ny", [
    ("imager_model", (0.0, 1.0, 0.0), 50, 50),
    ("small_wind", (0.3, -0.5, 0.8), 16, 12),
    ("wind32", (0.2, 0.7, -0.4), 24, 24),
])
def test_optical_depth_image_fp64_parity(which, raydir, nx, ny, request):
    from magritte_amd import solver
    from oracle import oracle

    m = request.getfixturevalue(which)
    op = m.options()
    fq = _image_case(m)
    print(f"[optical_depth_new] model={which} dir={raydir} {nx}x{ny}", flush=True)
    t, X, Y = solver.image_optical_depth_new(_dev(m), op, raydir, nx, ny, fq)
    t, X, Y = t.cpu().numpy(), X.cpu().numpy(), Y.cpu().numpy()
    tr, Xr, Yr = oracle.image_optical_depth_new(m, op, raydir, nx, ny, fq)
    np.testing.assert_array_equal(X.view(np.uint64), Xr.view(np.uint64))
    np.testing.assert_array_equal(Y.view(np.uint64), Yr.view(np.uint64))
    np.testing.assert_array_equal(t == 0.0, tr == 0.0)
    err = _err(t, tr)
    print(f"  max rel err {err:.3e}", flush=True)
    assert err < RTOL


@pytest.mark.parametrize("optical_depth", [False, True])
def test_image_ray_odd_point_count(optical_depth):
    """With at most 32 image frequencies the old imager takes two points per wave;
    51 points leave the last wave one, and every point must match the oracle."""
    import os

    from magritte_amd import models, solver
    from oracle import oracle

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    ld = models.read_lamda(os.path.join(root, "tests", "golden", "lamda_test.txt"), num=0)
    m = models.model_velocity_gradient_1d(ld, npoints=51, nrays=4, nquads=10,
                                          max_width_fraction=0.5)
    rr = m.hnrays - 1
    fq = _image_case(m)
    print(f"[image_ray odd] P={m.npoints} F={fq['nu'].shape[1]} optical_depth={optical_depth}",
          flush=True)
    out, X, Y = solver.image_ray(_dev(m), m.options(), rr, fq, optical_depth=optical_depth)
    ref, Xr, Yr = oracle.image_ray(m, m.options(), rr, fq, optical_depth=optical_depth)
    np.testing.assert_array_equal(X.cpu().numpy().view(np.uint64), Xr.view(np.uint64))
    np.testing.assert_array_equal(Y.cpu().numpy().view(np.uint64), Yr.view(np.uint64))
    err = _err(out.cpu().numpy(), ref)
    print(f"  max rel err {err:.3e}", flush=True)
    assert err < RTOL


def test_image_ray_mixed_precision(small_wind):
    from magritte_amd import solver
    from oracle import oracle

    m = small_wind
    fq = _image_case(m)
    dm = _dev(m)
    for optical_depth in (False, True):
        got, _, _ = solver.image_ray(dm, m.options(precision=_abi.PREC_MIXED), 3, fq,
                                     optical_depth=optical_depth)
        ref, _, _ = oracle.image_ray(m, m.options(), 3, fq, optical_depth=optical_depth)
        err = _err(got.cpu().numpy(), ref)
        print(f"  mixed optical_depth={optical_depth}: {err:.3e}", flush=True)
        assert err < 1e-5


def test_image_ray_arguments(small_wind, vgrad_model):
    from magritte_amd import solver

    dm = _dev(small_wind)
    with pytest.raises(_abi.MgxError, match="One line"):
        solver.image_ray(dm, small_wind.options(approx=_abi.APPROX_ONE_LINE), 0)
    with pytest.raises(_abi.MgxError, match="out of range"):
        solver.image_ray(dm, small_wind.options(), small_wind.nrays)
    with pytest.raises(_abi.MgxError, match=r"\(0,1,0\)"):
        solver.image_ray(_dev(vgrad_model), vgrad_model.options(), 0)
