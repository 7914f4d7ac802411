"""Host-side behaviour of the drop-in API that needs no GPU: factory
validation and error classes/messages of the reference (decomp.py:112-122,
375-378, 427-429, 753-756), and rejection of foreign matvecs."""

import numpy as np
import pytest

from matfree_amd import decomp, funm, operators, prng, stochtrace


def test_tridiag_sym_bad_reortho_raises_valueerror():
    with pytest.raises(ValueError, match="unsupported"):
        decomp.tridiag_sym(3, reortho="partial")


def test_tridiag_sym_reortho_flags():
    # decomp.py:112-122: "full" and "none" are accepted, anything else raises
    assert isinstance(decomp.tridiag_sym(3, reortho="none"), decomp._TridiagSymNone)
    assert isinstance(decomp.tridiag_sym(3, reortho="full"), decomp._TridiagSymFull)
    with pytest.raises(ValueError, match="unsupported"):
        decomp.tridiag_sym(3, reortho="partial")


@pytest.mark.parametrize("reortho_wrong", [True, "full_with_sparsity", "None"])
def test_hessenberg_bad_flag_raises_typeerror(reortho_wrong):
    # tests/test_decomp/test_hessenberg.py:72-75
    with pytest.raises(TypeError, match="Unexpected input"):
        decomp.hessenberg(1, reortho=reortho_wrong)


@pytest.mark.parametrize("k", [-1, 14])
def test_num_matvecs_out_of_range_raises_exceeds(k):
    # tests/test_decomp/test_consistency.py:50-63 -- raised before any device use
    A = np.eye(13, dtype=np.float32)
    v0 = np.ones(13, dtype=np.float32)
    with pytest.raises(ValueError, match="exceeds"):
        decomp.tridiag_sym(k)(operators.dense_matvec, v0, A)
    with pytest.raises(ValueError, match="exceeds"):
        decomp.tridiag_sym(k, reortho="none")(operators.dense_matvec, v0, A)
    for reortho in ("full", "none"):
        with pytest.raises(ValueError, match="exceeds"):
            decomp.hessenberg(k, reortho=reortho)(operators.dense_matvec, v0, A)
    integrand = funm.monte_carlo_funm_sym_logdet(decomp.tridiag_sym(k))
    with pytest.raises(ValueError, match="exceeds"):
        integrand(operators.dense_matvec, v0, A)


def test_python_callable_matvec_is_rejected():
    A = np.eye(4, dtype=np.float32)
    with pytest.raises(TypeError, match="operators"):
        decomp.tridiag_sym(2)(lambda v: A @ v, np.ones(4, np.float32))


def test_integrand_aliases():
    assert funm.integrand_funm_sym_logdet is funm.monte_carlo_funm_sym_logdet
    integrand = funm.integrand_funm_sym_logdet(decomp.tridiag_sym(4))
    assert integrand.num_matvecs == 4


def test_non_log_dense_funm_is_out_of_scope():
    with pytest.raises(NotImplementedError):
        funm.monte_carlo_funm_sym(funm.dense_funm_sym_eigh(np.exp), decomp.tridiag_sym(4))


def test_estimator_rejects_foreign_integrands():
    with pytest.raises(TypeError):
        stochtrace.estimator_monte_carlo(lambda m, v: 0.0, stochtrace.sampler_signs(np.ones(3), num=2))


def test_sampler_shapes_and_complex_rejected():
    s = stochtrace.sampler_signs({"a": np.ones(3), "b": np.ones((2, 2))}, num=5)
    assert s.n == 7 and s.num == 5
    with pytest.raises(NotImplementedError):
        stochtrace.sampler_signs(np.ones(3, dtype=complex), num=2)


def test_prng_key_and_words():
    assert prng.PRNGKey(1).tolist() == [0, 1]
    assert prng.key_words(prng.PRNGKey(2**33 + 7)) == (2, 7)
    with pytest.raises(ValueError):
        prng.key_words(np.zeros(3, np.uint32))


def test_rbf_operator_validation():
    with pytest.raises(ValueError):
        operators.RBF(np.zeros((4, 5)))
    with pytest.raises(ValueError):
        operators.RBF(np.zeros((4, 3)), precision="fp64")


# -- low_rank / linalg / gp (low_rank.py:56-198, tutorials/8_gaussian_logpdf.py) --


@pytest.mark.parametrize("factory,word", [("cholesky_partial", "n"),
                                          ("cholesky_partial_pivot", "nrows")])
def test_cholesky_rank_out_of_range_raises(factory, word):
    # low_rank.py:74-79 and 118-123: raised before any device use
    from matfree_amd import low_rank

    f = getattr(low_rank, factory)
    with pytest.raises(ValueError, match=f"Rank exceeds {word}: 6 >= 5."):
        f(operators.dense_matvec, nrows=5, rank=6)(np.eye(5))
    with pytest.raises(ValueError, match="Rank must be positive, but 0 < 1."):
        f(operators.dense_matvec, nrows=5, rank=0)(np.eye(5))


def test_cholesky_rejects_python_element_functions():
    from matfree_amd import low_rank

    cov = np.eye(4)
    with pytest.raises(TypeError, match="operators"):
        low_rank.cholesky_partial_pivot(lambda i, j: cov[i, j], nrows=4, rank=2)()


def test_cg_rejects_foreign_preconditioners():
    from matfree_amd import linalg

    with pytest.raises(TypeError):
        linalg.cg(lambda v: v, np.ones(3))


def test_serving_admission_rule():
    # HBM-footprint admission of the serving loop (matfree_amd.serving): a
    # request starts if it fits beside the running requests and what the idle
    # slots hold, after releasing idle slots (largest first) if needed; it
    # waits if it still does not fit while anything runs, and runs alone
    # otherwise
    from matfree_amd import serving

    adm = serving.admission
    held = {0: 0, 1: 0, 2: 0}
    assert adm(100, 0, held, [0, 1, 2], 0, False, None) == (True, [])
    # the server passes a list indexed by slot
    assert adm(100, 0, [0, 0, 0], [0, 1], 60, True, 150) == (False, [])
    assert adm(10, 0, [120, 0], [0, 1], 50, True, 150) == (True, [0])
    assert adm(100, 0, held, [0, 1, 2], 0, False, 150) == (True, [])
    assert adm(100, 0, held, [0, 1], 60, True, 150) == (False, [])  # 160 > 150: wait
    assert adm(100, 0, held, [0, 1], 50, True, 150) == (True, [])  # fits exactly
    assert adm(500, 0, held, [0, 1, 2], 0, False, 150) == (True, [])  # too big: alone
    # a slot that holds scratch is charged its holding plus the growth of its
    # pool (without the growth: the bound need); here that does not fit, so
    # the slot's own scratch is released
    assert adm(10, 0, {0: 120, 1: 0}, [0, 1], 50, True, 150) == (True, [0])
    assert adm(10, 0, {0: 90, 1: 0}, [0, 1], 50, True, 150) == (True, [])
    assert adm(10, 0, {0: 90, 1: 0}, [0, 1], 50, True, 140, 0) == (True, [])
    assert adm(10, 0, {0: 90, 1: 0}, [0, 1], 50, True, 140) == (True, [0])
    # idle slots give their memory back, largest first, only as far as needed
    held = {0: 0, 1: 80, 2: 40, 3: 10}
    assert adm(50, 0, held, [0, 1, 2, 3], 20, True, 150) == (True, [1])
    assert adm(100, 0, held, [0, 1, 2, 3], 20, True, 150) == (True, [1, 2])
    # when releasing everything is still not enough, nothing is released: wait
    assert adm(200, 0, held, [0, 1, 2, 3], 20, True, 150) == (False, [])
    # the chosen slot's own holding is released only when its holding plus
    # the growth exceeds what the request needs from an empty pool
    assert adm(100, 1, held, [0, 1, 2, 3], 0, False, 120, 20) == (True, [2])
    assert adm(100, 1, held, [0, 1, 2, 3], 0, False, 120) == (True, [1, 2])
    assert adm(30, 1, held, [0, 1, 2, 3], 0, False, 90) == (True, [1])
    assert adm(30, 1, held, [0, 1, 2, 3], 0, False, 90, 0) == (True, [1])
