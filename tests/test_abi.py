"""The C-ABI library loads and exports exactly what include/spca.h declares, and its
host-only helpers (settings defaults, reference clamp, formulation predicates) follow
the reference.  No device call is made: these run without a GPU."""
import ctypes
import os
import re

import pytest

import spca_mi355x as S

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "spca.h")


def _declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(spca_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = S.lib()
    declared = _declared()
    assert declared, "no declarations parsed"
    missing = [n for n in declared if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(S.EXPORTS) == declared


def test_version_string():
    assert "gfx950" in S.version()


def test_settings_defaults_match_reference():
    # OptimizationSettings() defaults, c++/src/class/optimization_settings.h:101-123
    c = S._Settings()
    S.lib().spca_settings_default(ctypes.byref(c))
    assert (c.tolerance, c.constraint, c.penalty, c.total_sp, c.batch, c.max_iter, c.use_otf) == \
        (0.01, 10, 0.0, 64, 64, 20, 0)
    assert c.formulation == S.SPCA_Formulation.L0_constrained_L2_PCA
    assert (c.sparse_products, c.exact_sign_products, c.exact_ties) == (1, 1, 1)


@pytest.mark.parametrize("s,n,total,batch,exp", [
    (10, 5, 64, 64, (5, 64, 64, 1)),      # s clamped to n
    (3, 100, 100, 32, (3, 128, 32, 4)),   # total rounded up to a batch multiple
    (3, 100, 100, 0, (3, 100, 100, 1)),   # batch 0 -> total
    (3, 100, 1, 7, (3, 7, 7, 1)),
])
def test_reference_clamp(s, n, total, batch, exp):
    # chceckInputAndModifyIt, c++/src/class/optimization_settings.h:156-168
    st = S.OptimizationSettings(constraintParameter=s, totalStartingPoints=total, batchSize=batch)
    tb = st.chceckInputAndModifyIt(n)
    assert (st.constraintParameter, st.totalStartingPoints, st.batchSize, tb) == exp


def test_formulation_predicates():
    # isConstrainedProblem / isL1ConstrainedProblem / isL1PenalizedProblem, :125-153
    F = S.SPCA_Formulation
    L = S.lib()
    cons = {f for f in F if L.spca_is_constrained(int(f))}
    assert cons == {F.L0_constrained_L2_PCA, F.L0_constrained_L1_PCA, F.L1_constrained_L2_PCA,
                    F.L1_constrained_L1_PCA}
    # the reference predicates are negations, so they are true outside their family too
    assert {f for f in F if not L.spca_is_l1_constrained(int(f))} == {F.L0_constrained_L1_PCA,
                                                                      F.L0_constrained_L2_PCA}
    assert {f for f in F if not L.spca_is_l1_penalized(int(f))} == {F.L0_penalized_L1_PCA, F.L0_penalized_L2_PCA}
    assert {f for f in F if L.spca_is_l1_variance(int(f))} == {F.L0_constrained_L1_PCA, F.L1_constrained_L1_PCA,
                                                               F.L0_penalized_L1_PCA, F.L1_penalized_L1_PCA}


def test_null_arguments_are_rejected_without_a_device():
    L = S.lib()
    assert L.spca_ctx_destroy(None) == S.SPCA_EINVAL
    assert L.spca_solve(None, 0, None, None, None, None, None, None) == S.SPCA_EINVAL
    assert L.spca_last_error(None) == b"null context"


@pytest.mark.parametrize("f", range(8))
@pytest.mark.parametrize("prec", [S.PREC_FP32, S.PREC_BF16X3, S.PREC_FP16])
def test_product_forms_and_the_fixed_point_row_limit(f, prec):
    """A^T*Z of the L1-variance formulations (z = sgn(A x)) runs the exact fixed-point product in
    FP32 and bf16x3 while its int32 sums cannot overflow (padded rows < 2^18); taller matrices,
    fp16 and the L2-variance formulations run the precision's own product."""
    st = S.OptimizationSettings(formulation=f, precision=prec)
    l1v = f in (1, 3, 5, 7)
    exact = l1v and prec in (S.PREC_FP32, S.PREC_BF16X3)
    assert st.product_forms(4096, 65536) == (prec, S.FORM_FIX8 if exact else prec)
    # 261888 rows pad to 2^18 - 256; one more row pads to 2^18, past the accumulators' bound
    assert st.product_forms(261888, 4096)[1] == (S.FORM_FIX8 if exact else prec)
    assert st.product_forms(261889, 4096)[1] == prec
    st.exact_sign_products = 0
    assert st.product_forms(4096, 65536) == (prec, prec)


def test_product_forms_rejects_what_a_solve_rejects():
    st = S.OptimizationSettings()
    assert st.product_forms(100, 2 ** 23 - 1) == (0, 0)
    with pytest.raises(S.SpcaError) as e:
        st.product_forms(100, 2 ** 23)  # the step kernel's fixed-point sums need n < 2^23
    assert e.value.code == S.SPCA_EINVAL
    L = S.lib()
    c = st._c()
    assert L.spca_product_forms(None, 10, 10, None, None) == S.SPCA_EINVAL
    assert L.spca_product_forms(ctypes.byref(c), 0, 10, None, None) == S.SPCA_EINVAL
    c.formulation = 8
    assert L.spca_product_forms(ctypes.byref(c), 10, 10, None, None) == S.SPCA_EINVAL


def test_double_precision_settings_are_rejected():
    """The reference's useDoublePrecision (optimization_settings.h:71-99) selects a double solve,
    which this solver does not have: the settings never reach the library as an FP32 solve."""
    st = S.OptimizationSettings(useDoublePrecision=True)
    with pytest.raises(S.SpcaError) as e:
        st._c()
    assert e.value.code == S.SPCA_EINVAL
