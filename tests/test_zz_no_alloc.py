"""qsm_r2r3d_execute_host_noalloc: the allocation-free form of
qsm_r2r3d_execute_host, through the C ABI, against that entry point on the
same inputs (the same pass kernels, so the outputs are equal bit for bit),
on two calls in a row, the second with a smaller field, in both precisions;
and the size limit of its staging buffer."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", (0, 1))
def test_noalloc_matches_execute_host(qsm_mod, prec):
    rng = np.random.default_rng(21 + prec)
    for shape, kinds in (((64, 48, 96), (9, 5, 5)), ((32, 16, 24), (4, 8, 4))):
        print(f"noalloc prec {prec} shape {shape} kinds {kinds}", flush=True)
        plan = qsm_mod.Plan(shape, kinds, prec)
        x = rng.standard_normal(shape).astype(plan.dtype)
        want = plan(x)
        got = plan.noalloc(x)
        assert np.array_equal(got, want), (shape, float(np.max(np.abs(got - want))))
        plan.close()


def test_noalloc_rejects_fields_beyond_the_buffer(qsm_mod):
    plan = qsm_mod.Plan((256, 256, 512), (5, 5, 5), qsm_mod.FP64)  # 256 MiB > 128 MiB
    x = np.zeros((256, 256, 512))
    with pytest.raises(qsm_mod.QsmError) as e:
        plan.noalloc(x)
    assert e.value.code == -2 and "staging buffer" in str(e.value)
