"""Tutorial 03 of the reference (examples/03_seq_file.py): a CW saturation written as a Pulseq file, read
back, converted and simulated, against the same saturation built in code. The file route steps the 2 s
block pulse in 20 steps of 0.1 s; the propagators compose exactly, so in float64 the spectra agree to
rounding."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

B0_T = 2.8936


def _params(f=0.005, k=300.0, dw=3.5):
    from bmc_jax.simulation import SimParams
    s = np.float64
    return SimParams(water_r1=s(1 / 3.0), water_r2=s(1 / 0.6), water_f=s(1.0), cest_k=s([k]), cest_f=s([f]),
                     cest_dw=s([dw]), cest_r1=s([1.0]), cest_r2=s([100.0]), mt_k=s(0), mt_f=s(0), mt_dw=s(0),
                     mt_r1=s(0), mt_r2=s(1), mt_lineshape=np.int32(0), b0=s(B0_T), gamma=s(267.5153),
                     b0_inhom=s(0), rel_b1=s(1.0))


@pytest.fixture()
def x64():
    import bmc_jax
    from bmc_jax import _native
    _native.torch_cuda()
    was = bmc_jax.settings.x64_enabled()
    bmc_jax.settings.use_x64(True)
    yield
    bmc_jax.settings.use_x64(was)


def test_cw_seq_file_simulates_like_code(x64, tmp_path):
    from bmc_jax.sequences import cw_saturation, hz_per_ppm
    from bmc_jax.sequences.pulseq import SeqWriter, events_from_seq_batch, read_seq
    from bmc_jax.simulation import initial_state, simulate_spectrum
    from bmc_jax.simulation.events import cast_events
    ppm = np.linspace(-6.0, 6.0, 25)
    b1, tsat = 1.0, 2.0
    code = cast_events(cw_saturation(ppm, b1, tsat, b0_T=B0_T), np.float64)
    w = SeqWriter()
    for p in ppm:
        w.block_pulse(float(code.rf_amp[0, 0]), tsat, freq_offset=p * hz_per_ppm(B0_T))
        w.adc()
    path = tmp_path / "saturation.seq"
    w.write(path)
    ev, n = events_from_seq_batch(read_seq(str(path)))
    assert n == ppm.size
    # the file stores the offsets in Hz with 17 significant digits; the code-built events round them to float32
    code = code._replace(rf_freq=np.where(code.type == 1, ppm[:, None] * hz_per_ppm(B0_T), 0.0))
    p = _params()
    z_file = np.asarray(simulate_spectrum(p, ev, initial_state(p), return_full_trajectory=False)[:, 4])
    z_code = np.asarray(simulate_spectrum(p, code, initial_state(p), return_full_trajectory=False)[:, 4])
    assert z_file.min() < 0.9
    np.testing.assert_allclose(z_file, z_code, atol=1e-11, rtol=0)
