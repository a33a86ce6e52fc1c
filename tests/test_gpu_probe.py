"""The measured BF16 matrix-core ceiling (km_measure_bf16_mfma_peak): a pure MFMA
loop reaches a rate below the published dense peak (2.5 PFLOP/s at the peak
clock) and above a third of it, and the shader clock it reports is plausible."""
import pytest

from onedal_amd.engine import InvalidArgument

pytestmark = pytest.mark.gpu


def test_ceiling_is_plausible(engine):
    tflops, ghz = engine.measure_bf16_mfma_peak(5000)
    print("measured BF16 MFMA ceiling %.1f TFLOP/s at %.3f GHz" % (tflops, ghz), flush=True)
    assert 800.0 < tflops < 2600.0
    assert 0.8 < ghz < 3.0


def test_ceiling_rejects_bad_iterations(engine):
    with pytest.raises(InvalidArgument):
        engine.measure_bf16_mfma_peak(0)
    with pytest.raises(InvalidArgument):
        engine.measure_bf16_mfma_peak(1 << 25)
