"""The grouped GEMM's kernel forms and tile shapes give bit-identical amplitudes.

Every contraction of the residual (ccm_t2_eqn.f90:431-1026) runs through one
grouped complex GEMM whose tile shape and kernel form the autotuner picks per
tile class: B read from L2, B staged in LDS, or A and B staged in LDS, with
64..256-row and 16..80-column tiles.  All of them sum k in the same order and
split k by the group alone, so the amplitudes must not depend on the choice.
Each case forces one form on every class of every family (CCGPU_SHAPES, which
also switches the autotuner off) and compares T2 after a few iterations with
the first form's, element for element.  The pp ladder family has no split-K
and, for N <= 4, one form of its own (the LDS-pipelined MFMA kernel); the
pipelined form (lb 3: V and T streamed into LDS by direct-to-LDS loads) is
forced on the pp ladder alone, the other families keeping B from L2."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# mb.nb.lb: 64*mb rows x 16*nb columns; lb 0 = B from L2, 1 = B in LDS, 2 = A and B in LDS,
# 3 = pp ladder through the LDS load pipeline, 4 = the same with eight waves (64 x 32 tiles)
FORMS = ["1.1.0", "1.1.1", "1.1.2", "1.2.2", "1.3.2", "2.1.0", "2.2.1", "4.1.0", "1.5.1", "1.1.3", "1.2.3",
         "1.2.4"]
TAGS = ("", "L", "A", "P", "W")


def t2_after(ccnm, monkeypatch, form, precision):
    other = form if not (form.endswith(".3") or form.endswith(".4")) else "1.1.0"
    spec = ",".join(f"{f}.{c}={form if f == 1 else other}" for f in range(5) for c in range(5))
    monkeypatch.setenv("CCGPU_SHAPES", spec)
    monkeypatch.setenv("CCGPU_TIMING", "2")   # names every launched class with its shape
    s = ccnm.Solver(A=66, matter=ccnm.PNM, rho=0.16, nmax=3)
    try:
        s.generate_interaction(ccnm.INT_MINNESOTA)
        s.setup_t_amplitudes()
        r = s.ccd_iter(subspace=8, diis_step=8, tol=0.0, maxit=4, precision=precision)
        assert r["iterations"] == 4
        # the forced form ran: every MFMA class of the five families carries it
        def tag(f):
            mb, nb, lb = (int(x) for x in f.split("."))
            return f"{mb}x{nb}" + TAGS[lb]
        launched = [k for k in ccnm.kernel_stats() if k.startswith("fam") and ".c5." not in k]
        assert {k.split(".")[0] for k in launched} == {f"fam{f}" for f in range(5)}, launched
        for k in launched:
            want = tag(form) if k.startswith("fam1.") else tag(other)
            assert k.endswith("." + want), (want, launched)
        return s.t2_unique()
    finally:
        s.close()
        monkeypatch.delenv("CCGPU_SHAPES")
        monkeypatch.delenv("CCGPU_TIMING")


@pytest.mark.parametrize("precision", [0, 1, 2], ids=["fp64", "fp32", "v32"])
def test_gemm_forms_bit_identical(ccnm, monkeypatch, precision):
    ref = t2_after(ccnm, monkeypatch, FORMS[0], precision)
    assert np.all(np.isfinite(ref)) and np.abs(ref).max() > 0
    for form in FORMS[1:]:
        t2 = t2_after(ccnm, monkeypatch, form, precision)
        diff = np.abs(t2 - ref).max()
        assert np.array_equal(t2, ref), (form, diff)
