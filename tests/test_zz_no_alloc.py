"""GPU test of qvti_fft3d_noalloc, the allocation-free variant of qvti_fft3d:
the same transform on caller-owned device data, twiddle tables in the
library's static per-device scratch array.  Checked bit for bit against
qvti_fft3d (host data, library-allocated buffers) on two calls in a row, the
second on a smaller grid (the scratch keeps the first call's tables, which
the second must rebuild), and in both directions."""

import numpy as np
import pytest

import hiprt
import qvti

pytestmark = pytest.mark.gpu


def _noalloc(x, sign, stream):
    """The transform through qvti_fft3d_noalloc on a device buffer the test
    allocates (hipMalloc of the library's own HIP runtime)."""
    nx, ny, nz = x.shape
    x = np.ascontiguousarray(x, np.complex64)
    out = np.empty_like(x)
    with hiprt.DeviceBuffer(x.nbytes) as d:
        d.upload(x)
        print("qvti_fft3d_noalloc %dx%dx%d sign %d" % (nx, ny, nz, sign), flush=True)
        qvti.check(qvti.lib().qvti_fft3d_noalloc(nx, ny, nz, sign, d.ptr, stream.s))
        stream.synchronize()
        d.download(out)
    return out


@pytest.mark.parametrize("sign", [-1, 1])
def test_noalloc_fft3d_equals_fft3d_on_two_calls(sign):
    rng = np.random.default_rng(3)
    st = hiprt.Stream()
    try:
        for shape in [(64, 48, 40), (20, 16, 24)]:
            x = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
            want = qvti.fft3d(x, sign)
            got = _noalloc(x, sign, st)
            assert np.array_equal(got, want), (shape, sign, float(np.abs(got - want).max()))
            ref = np.fft.fftn(x.astype(np.complex128)) if sign < 0 else np.fft.ifftn(x.astype(np.complex128)) * x.size
            assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-6
    finally:
        st.close()


def test_noalloc_fft3d_rejects_axes_beyond_the_scratch():
    with hiprt.DeviceBuffer(2 * 4 * 16 * 16 * 1024) as d:
        with pytest.raises(qvti.QvtiError) as err:
            qvti.check(qvti.lib().qvti_fft3d_noalloc(16, 16, 1024, -1, d.ptr, None))
    assert err.value.code == qvti.E_UNSUPPORTED
    assert "scratch" in str(err.value)
