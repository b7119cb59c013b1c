"""Measurement probes: the VALU issue-rate kernel gives rates a gfx950 chip
can issue (wave64 v_fma_f32 at most one per 4 cycles per SIMD, a
transcendental no faster than an FMA), the latent sampler probe a draw rate
that its instruction count allows, and the copy probe a bandwidth under the
HBM peak."""
import pytest

from bsglmm_amd import copy_bandwidth, sampler_rate, valu_rate

pytestmark = pytest.mark.gpu

NOMINAL = 1024 * 2.4e9 / 4  # 1024 SIMDs x 2.4 GHz / 4 cycles per wave64 instruction


def test_valu_rate_is_plausible():
    fma, ms = valu_rate("fma", 5)
    trans, ms_t = valu_rate("trans", 2)
    assert 0.4 * NOMINAL < fma < 1.05 * NOMINAL, fma
    assert 0.1 * NOMINAL < trans <= 1.05 * fma, (trans, fma)
    assert ms > 0.0 and ms_t > 0.0


def test_sampler_rate_is_plausible():
    """The sampler alone issues ~50 VALU instructions per draw (5 of them
    half-rate transcendentals), so its rate lies between the VALU issue rate
    times 64 lanes over 100 and over 25 instructions per draw."""
    dps, ms = sampler_rate(3)
    assert NOMINAL * 64 / 100 < dps < NOMINAL * 1.2 * 64 / 25, dps
    assert ms > 0.0


def test_copy_bandwidth_is_plausible():
    gbs, ms = copy_bandwidth(1 << 28, 5)
    assert 500.0 < gbs < 8000.0 and ms > 0.0
