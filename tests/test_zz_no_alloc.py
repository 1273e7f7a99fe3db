"""bk_learner_root_histogram_noalloc (the entry point that allocates nothing: its
partial histograms live in the library's file-scope device scratch array)
against bk_learner_root_histogram on the same learner configuration, through the
C ABI: two calls in a row on one device, the second on a smaller dataset, so
the second call reuses the scratch the first one left behind.

Bar: quantized histograms and every count bit for bit; the float forms within
the rounding of two summation orders of the same f32 inputs
(4 * n_b * 2^-53 * sum|x| per bin, twice the bound each order keeps from the
sequential sum). A configuration whose partials exceed the scratch is refused
with an error that names the limit.
"""
import sys

import numpy as np
import pytest

import bankai_gpu as bg
import synth

pytestmark = pytest.mark.gpu


def announce(msg):
    print(f"\n[case] {msg}", flush=True)
    sys.stdout.flush()


FORMS = {
    "quantized_b4_stochastic": dict(use_quantized_grad=1, num_grad_quant_bins=4, stochastic_rounding=1, seed=5,
                                    num_threads=3, form=bg.FORM_QUANTIZED),
    "quantized_b127_nearest": dict(use_quantized_grad=1, num_grad_quant_bins=127, stochastic_rounding=0,
                                   form=bg.FORM_QUANTIZED),
    "plain_f64": dict(form=bg.FORM_PLAIN_F64),
    "f32_compensated": dict(form=bg.FORM_F32_COMPENSATED),
}


def abs_bin_sums(ds, x):
    ax = np.abs(x.astype(np.float64))
    return np.concatenate([np.bincount(c, weights=ax, minlength=int(nb))
                           for c, nb in zip(ds.group_columns(), ds.group_num_total_bin())])


def one_call(rows, seed, form, const_hess):
    import torch
    x, _ = synth.binary_28(rows, seed=seed)
    ds = bg.Dataset(x)
    rng = np.random.default_rng(seed)
    g = (rng.standard_normal(ds.num_data) * np.exp(rng.uniform(-4, 4, ds.num_data))).astype(np.float32)
    h = np.full(ds.num_data, 0.5, np.float32) if const_hess else rng.uniform(0.01, 0.3, ds.num_data).astype(np.float32)
    p = bg.LearnerParams(num_leaves=7, **FORMS[form])
    # two learners of one configuration: each call is its learner's first tree, so
    # a stochastic learner draws the same noise rotation in both
    want_h, want_c = bg.Learner(ds, p, const_hess).root_histogram(g, h)
    learner = bg.Learner(ds, p, const_hess)
    dev = torch.device("cuda:0")
    tg, th = torch.as_tensor(g, device=dev), torch.as_tensor(h, device=dev)
    tb = want_c.shape[0]
    out = torch.full((tb, 2), -7, dtype=torch.int64 if p.form == bg.FORM_QUANTIZED else torch.float64, device=dev)
    cnt = torch.full((tb,), -7, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    learner.root_histogram_noalloc(tg.data_ptr(), th.data_ptr(), out.data_ptr(), cnt.data_ptr())
    got_h, got_c = out.cpu().numpy(), cnt.cpu().numpy()
    learner.close()
    assert np.array_equal(got_c, want_c)
    if p.form == bg.FORM_QUANTIZED:
        assert np.array_equal(got_h, want_h)
    else:
        n_b = want_c.astype(np.float64)
        tol_g = 4.0 * n_b * 2.0 ** -53 * abs_bin_sums(ds, g)
        assert np.all(np.abs(got_h[:, 0] - want_h[:, 0]) <= tol_g)
        if const_hess:
            assert np.array_equal(got_h[:, 1], want_h[:, 1])
        else:
            tol_h = 4.0 * n_b * 2.0 ** -53 * abs_bin_sums(ds, h)
            assert np.all(np.abs(got_h[:, 1] - want_h[:, 1]) <= tol_h)


@pytest.mark.parametrize("const_hess", [0, 1])
@pytest.mark.parametrize("form", sorted(FORMS))
def test_noalloc_root_histogram_matches_allocating_entry(form, const_hess):
    announce(f"root histogram without allocation, {form}, const_hess={const_hess}: 300k then 40k rows")
    one_call(300_000, 31, form, const_hess)
    one_call(40_000, 32, form, const_hess)


def test_noalloc_refuses_partials_beyond_the_scratch():
    import torch
    announce("root histogram without allocation: partials beyond the scratch are refused")
    rng = np.random.default_rng(2)
    x = rng.random((170_000, 300), dtype=np.float32)
    ds = bg.Dataset(x)
    p = bg.LearnerParams(num_leaves=7, use_quantized_grad=1, num_grad_quant_bins=127, form=bg.FORM_QUANTIZED)
    learner = bg.Learner(ds, p, False)
    dev = torch.device("cuda:0")
    g = torch.ones(ds.num_data, dtype=torch.float32, device=dev)
    tb = int(np.sum(ds.group_num_total_bin()))
    out = torch.zeros((tb, 2), dtype=torch.int64, device=dev)
    cnt = torch.zeros(tb, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    with pytest.raises(bg.BankaiError, match="MiB scratch"):
        learner.root_histogram_noalloc(g.data_ptr(), g.data_ptr(), out.data_ptr(), cnt.data_ptr())
    learner.close()
