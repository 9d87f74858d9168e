"""The no-allocation power method, pwm_power_method_device_noalloc, against
the entry point it varies, pwm_power_method_device, through the C ABI.

Both run the same kernels; the variant takes its workspace (step control,
step records, FP64 partial sums) from the library's fixed per-GPU array
instead of buffers the handle allocates.  Bars: the CRS and merge-path
kernels sum in a fixed order, so their results are bit-identical; CSB and
BlockCOH add into LDS with atomics (order varies), so theirs agree within
1e-12 (vectors) and 1e-10 (eigenvalue).  Two calls in a row, the second on a
smaller matrix with fewer steps; neither allocates device memory."""
import ctypes

import numpy as np
import pytest

import oracle

pw = pytest.importorskip("thesis_spmv_amd")
pytestmark = pytest.mark.gpu

FORMATS = [("crs_tiles", pw.CRSHip, {"crs_kernel": 2}, True), ("crs_vector", pw.CRSHip, {"crs_kernel": 1}, True),
           ("merge_path", pw.CRSMergeHip, {}, True), ("csb", pw.CSBHip, {}, False),
           ("csb_swept", pw.CSBHip, {"csb_layout": 4}, False),
           ("blockcoh", pw.BlockCOHHip, {}, False)]


def _hip_runtime():
    pw.lib()
    with open("/proc/self/maps") as f:
        paths = {ln.split()[-1] for ln in f if "libamdhip64.so" in ln}
    assert paths, "libamdhip64 not loaded"
    return ctypes.CDLL(sorted(paths)[0])


def _free_device_bytes(hip):
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def _load(cls, A, opt):
    m = cls(-1, "f64", **opt)
    m.loadCRS(A.nor, A.noc, A.row_start, A.col_ind, A.data)
    return m


def _run(m, vecs, x0, its, tol, noalloc):
    x, y = vecs
    x.upload(x0)
    y.fill(0.0)
    fn = m.power_method_device_noalloc if noalloc else m.power_method_device
    lam, yn, k = fn(x, y, its, tol)
    return lam, yn, k, x.download(), y.download()


@pytest.mark.parametrize("name,cls,opt,exact", FORMATS, ids=[f[0] for f in FORMATS])
def test_noalloc_power_method_matches_the_allocating_one(name, cls, opt, exact):
    hip = _hip_runtime()
    rs, ci, dv = pw.gen_spd_powerlaw(30000, seed=17)
    big = oracle.CRSMatrix(30000, 30000, rs, ci, dv)
    small = oracle.poisson_crs(40, 30)
    xb = pw.gen_uniform(big.nor, 1)
    xs = pw.gen_uniform(small.nor, 2)
    mb, ms = _load(cls, big, opt), _load(cls, small, opt)
    # a stop test on the fixed-order kernels; fixed step counts on the
    # atomic-order formats (a stop test could end them a step apart)
    tol = 1e-9 if exact else -1.0
    vb = (pw.DeviceVector(big.nor), pw.DeviceVector(big.nor))
    vs = (pw.DeviceVector(small.nor), pw.DeviceVector(small.nor))
    bytes_b, bytes_s = mb.info()["device_bytes"], ms.info()["device_bytes"]
    free0 = _free_device_bytes(hip)
    got_b = _run(mb, vb, xb, 400, tol, True)
    got_s = _run(ms, vs, xs, 37, -1.0, True)  # second call: smaller matrix, fewer steps
    free1 = _free_device_bytes(hip)
    assert mb.info()["device_bytes"] == bytes_b and ms.info()["device_bytes"] == bytes_s
    assert free1 == free0, (free0, free1)
    ref_b = _run(mb, vb, xb, 400, tol, False)
    ref_s = _run(ms, vs, xs, 37, -1.0, False)
    assert mb.info()["device_bytes"] > bytes_b  # the allocating entry point allocated its workspace
    for got, ref, what in ((got_b, ref_b, "power-law SPD 30000"), (got_s, ref_s, "Poisson 40x30")):
        lam, yn, k, x, y = got
        lr, ynr, kr, xr, yr = ref
        print(f"{name} {what}: steps {k} / {kr}, lambda {lam!r} / {lr!r}", flush=True)
        assert k == kr
        if exact:
            assert lam == lr and yn == ynr and np.array_equal(x, xr) and np.array_equal(y, yr)
        else:
            assert abs(lam - lr) <= 1e-10 * abs(lr) and abs(yn - ynr) <= 1e-10 * abs(ynr)
            assert np.max(np.abs(x - xr)) <= 1e-12 and np.max(np.abs(y - yr)) <= 1e-12
    # against the oracle too: the smaller call's eigenvalue estimate
    xo, yo = oracle.crs_power_iteration(small, xs, 36)
    v = xo if 36 % 2 == 0 else yo
    ref = float(v @ oracle.crs_mv(small, v) / (v @ v))
    assert abs(got_s[0] - ref) <= 1e-10 * abs(ref)


def test_noalloc_limits_are_refused_before_any_launch():
    A = oracle.poisson_crs(20, 20)
    m = _load(pw.CRSMergeHip, A, {})
    x = pw.DeviceVector.from_array(np.ones(A.nor))
    y = pw.DeviceVector.from_array(np.full(A.nor, 3.0))
    with pytest.raises(pw.PwmError) as e:
        m.power_method_device_noalloc(x, y, 65537, -1.0)
    assert e.value.code == -1
    assert np.all(y.download() == 3.0) and np.all(x.download() == 1.0)
    lam, _, k = m.power_method_device_noalloc(x, y, 65536 // 1024, -1.0)  # within the limit
    assert k == 64 and lam > 0
