"""The conditioning queue (the reference's --queue, fp_filter.py:300-305): datasets conditioned on the queue's thread
and stream while the previous ones are filtered give the same triggers as conditioning and filtering one after the
other, over several rounds (the queue's tensors are handed to the filtering stream, not reused early)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_queue_equals_serial(gpu):
    import fp_filter, synth, workloads
    sr = 2048.0
    bank = synth.circular_bank(64, seed=11, mmin=5, mmax=100)
    inj = [synth.Injection(30.0, float(bank.m1[5]), float(bank.m2[5]), 14.0, 0.3)]
    streams = [synth.make_stream('H1', 128, sr, seed=1, injections=inj, n_glitch=2, glitch_snr=300.0),
               synth.make_stream('L1', 128, sr, seed=2, injections=inj)]
    cond = workloads.Conditioning(128, 16, 16, 8, 4, 29, 8, autogating_threshold=50.0)
    x_dev = [torch.from_numpy(st.x).cuda() for st in streams]

    def condition():
        return workloads.condition_streams(streams, cond, [x.clone() for x in x_dev])

    opts = fp_filter.Options(tile=32, precision='bf16', snr_threshold=5.0)
    ds0 = condition()
    IPall, sw, isw = fp_filter.svd_metric(ds0)
    plan = fp_filter.BankPlan(bank.m1, bank.m2, None, ds0[0].flen, ds0[0].df, opts, sw, isw)

    def run(ds):
        return [t.to_dict(gates=d.gates) for t, d in zip(fp_filter.Filter(ds, opts).run(plan), ds)]

    ref = run(condition())
    assert sum(len(o['snr']) for o in ref) > 0 and any(len(d.gates) for d in ds0)
    q = workloads.PrepQueue()
    q.submit(condition)
    for k in range(4):
        ds = q.get()
        if k < 3:
            q.submit(condition)
        got = run(ds)
        for a, b in zip(ref, got):
            for key in a:
                np.testing.assert_array_equal(a[key], b[key], err_msg=f'round {k} {key}')
