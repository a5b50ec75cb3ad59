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
