"""Device peak probes (probe_kernels.hip): plausible magnitudes on an MI355X-class device."""
import pytest

pytestmark = pytest.mark.gpu


def test_device_peaks_plausible(ip):
    print("device peaks", flush=True)
    p = ip.device_peaks()
    # published MI355X: 8 TB/s HBM, 157.3 TFLOP/s FP32 and 78.6 TFLOP/s FP64 matrix
    assert 1000 < p["copy_gbs"] < 9000
    assert 20 < p["mfma_f32_tflops"] < 170
    assert 10 < p["mfma_f64_tflops"] < 90
