"""The reference's published validation (measured reflection against Fresnel,
MasterController.py:533-569, BaseFDTD11.py:882-923) on the device: a
Lorentz half-space measured by the two-pass protocol through the batched
sweep path (sweep.run_alone, device DFT at 256 frequencies) equals the
closed-form reflection of the interface on the Yee grid
(refcases.discrete_fresnel) to 2e-6 in FP64 and 1e-4 in FP32 wherever the
pulse carries power (246 of 256 frequencies), and the continuum Fresnel of
the analytic Lorentz eps to 2e-2."""
import numpy as np
import pytest

from refcases import discrete_fresnel, fresnel_spec

pytestmark = pytest.mark.gpu

FREQS = np.linspace(100e12, 300e12, 256)


@pytest.mark.parametrize("prec,tol", [("f64", 2e-6), ("f32", 1e-4)])
def test_device_half_space_reflection_is_discrete_fresnel(prec, tol):
    from fdtd_pic_amd import sweep
    p = fresnel_spec()
    job = sweep.run_alone(p, prec, freqs=FREQS)
    band = np.abs(job.X[0]) > 0.1 * np.abs(job.X[0]).max()
    r_num, r_cont = discrete_fresnel(p, FREQS)
    err = np.abs(job.R - r_num)[band].max()
    print(f"{prec}: in-band {band.sum()}, max |R - r_grid| {err:.2e}, "
          f"max |R - r_Fresnel| {np.abs(job.R - r_cont)[band].max():.2e}", flush=True)
    assert band.sum() >= 240
    assert err < tol
    assert np.abs(job.R - r_cont)[band].max() < 2e-2
