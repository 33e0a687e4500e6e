"""GPU test of gbh_force_evidence_noalloc, the evidence entry point that
allocates nothing (its per-target evidence is staged in the library's per-GPU
scratch array): on two calls in a row, the second with a smaller tree, its
evidence and forces equal gbh_force_evidence's bit for bit; trees beyond the
scratch size are refused with an error."""
import numpy as np
import pytest

from galaxy_amd import BhError, Config, NbodyGpu, ParallelTreeRuntime, ics

pytestmark = pytest.mark.gpu


def run_both(gpu, rt, state, cfg):
    ev = gpu.create_evolving_state(state)
    tree = rt.create_tree(gpu, ev, large=len(state) > 65536)
    rt.rebuild(gpu, ev, tree, cfg)
    st_a, h_a = rt.force_evidence(ev, tree, cfg)
    acc_a = gpu.read_evolving_accelerations(ev)
    st_b, h_b = rt.force_evidence_noalloc(ev, tree, cfg)
    acc_b = gpu.read_evolving_accelerations(ev)
    tree.close()
    ev.close()
    return (st_a, h_a, acc_a), (st_b, h_b, acc_b)


def test_noalloc_evidence_equals_allocating_entry_point():
    gpu = NbodyGpu()
    rt = ParallelTreeRuntime(gpu)
    cfg = Config(theta=0.5, softening_kpc=0.005, g=1.0, bucket=4, max_depth=16)
    # the largest tree the scratch holds, then a smaller one (the scratch still
    # holds the first call's evidence beyond the second tree's bodies)
    for n, seed in [(65536, 303), (9001, 7)]:
        print(f"noalloc evidence: n={n}", flush=True)
        a, b = run_both(gpu, rt, ics.uniform_collapse(n, seed), cfg)
        assert np.array_equal(a[0], b[0]), n
        assert np.array_equal(a[1], b[1]), n
        assert np.array_equal(a[2].view(np.uint32), b[2].view(np.uint32)), n
        assert int(a[0][:, 2].sum()) > 0
    print("noalloc evidence: refuses n=70000", flush=True)
    ev = gpu.create_evolving_state(ics.uniform_collapse(70000, 5))
    tree = rt.create_tree(gpu, ev, large=True)
    rt.rebuild(gpu, ev, tree, cfg)
    with pytest.raises(BhError, match="at most 65536 bodies"):
        rt.force_evidence_noalloc(ev, tree, cfg)
    tree.close()
    ev.close()
    rt.close()
    gpu.close()
