"""The allocation-free entry point w90gpu_omega_noalloc against w90gpu_omega
(the same kernels with buffers the library allocates): equal results on
two calls in a row, the second with a smaller problem, and a clear error
for a problem larger than the static scratch array.  Calls through the C
ABI (libw90gpu.so)."""
import numpy as np
import pytest

from conftest import Case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from wannier90_amd import capi

    c = capi.Context(0)
    yield c
    c.close()


def test_noalloc_omega_equals_omega(ctx):
    from wannier90_amd import synth

    # the fcc 16-band 512-k problem (64-way k-range reduction, the MFMA-free
    # 16-band class), then a 4-band 8-k fixture, then a 40-band problem (the
    # MFMA tile class), each twice: bit-identical to w90gpu_omega
    big = synth.config(2)
    _, m_big = synth.initial_state(big)
    small = Case("ref_basic1")
    mid = synth.make_problem(4242 + 40, kind="sc", nw=40, mp=(3, 3, 3), decay=2.0)
    _, m_mid = synth.initial_state(mid)
    cases = [("fcc16-512k", big.kmesh, m_big), ("ref_basic1", small.kmesh, small.m0),
             ("sc40-27k", mid.kmesh, m_mid), ("ref_basic1 again", small.kmesh, small.m0)]
    for name, km, m in cases:
        m = np.asfortranarray(m)
        for first, inv in ((True, 0.0), (False, 0.75)):
            print(f"[noalloc] {name} first_pass={first}", flush=True)
            ref = ctx.omega(km, m, first, inv)
            got = ctx.omega_noalloc(km, m, first, inv)
            for a, b in zip(ref, got):
                assert np.array_equal(a, b), name


def test_noalloc_rejects_problem_above_scratch(ctx):
    from wannier90_amd import capi, synth

    # 64 bands on 6x6x6 k with nntot 12: M alone is 3 x 127 MB
    p = synth.make_problem(11, kind="bcc", nw=64, mp=(6, 6, 6), decay=2.0, with_m=False)
    m = np.zeros((64, 64, p.kmesh.nntot, p.num_kpts), dtype=np.complex128, order="F")
    with pytest.raises(capi.W90Error) as e:
        ctx.omega_noalloc(p.kmesh, m)
    assert e.value.code == capi.ERR_INPUT and "scratch" in str(e.value)
