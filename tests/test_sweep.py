"""Host logic of wavenumber sweeps and of the shift model's data handling."""
import numpy as np

from tailor_amd import sweep as sw


def test_sweep_instances_are_deterministic_and_in_range():
    a = sw.sweep_instances(64, seed=1, sizes=[(100, 50), (200, 100)])
    b = sw.sweep_instances(64, seed=1, sizes=[(100, 50), (200, 100)])
    assert [i.alpha for i in a] == [i.alpha for i in b]
    al = np.array([i.alpha for i in a])
    assert np.all((al >= 0.5) & (al <= 3.0))
    assert sum(i.ny == 100 for i in a) == 32 and sum(i.ny == 200 for i in a) == 32
    assert len({i.seed for i in a}) == 64          # per-instance base-flow seeds
    c = sw.sweep_instances(8, seed=2, sizes=[(100, 50)])
    assert [i.alpha for i in c] != [i.alpha for i in a[:8]]


def test_training_instances_cover_the_box():
    t = sw.training_instances(500, seed=2)
    x = np.array([(i.alpha, i.reynolds, i.mach) for i in t])
    lo, hi = x.min(0), x.max(0)
    assert np.all(lo >= [0.5, 500, 0.3]) and np.all(hi <= [3.0, 3000, 0.9])
    assert np.all(hi - lo > [2.2, 2200, 0.5])


def test_alpha_order_groups_sizes_then_wavenumber():
    insts = sw.sweep_instances(16, seed=1, sizes=[(100, 50), (200, 100)])
    order = sw.alpha_order(insts)
    keys = [(insts[k].ny * insts[k].nz, insts[k].alpha) for k in order]
    assert keys == sorted(keys)


def test_sweep_result_totals():
    inst = sw.Instance(10, 8, 0, 1.0)
    recs = [sw.Record(inst, 0.1j, 8, 8, np.zeros(8), np.zeros(8), 2, 64, 0, 1, False, 0.5),
            sw.Record(inst, 0.1j, 8, 6, np.zeros(8), np.zeros(8), 3, 80, 0, 2, True, 0.7)]
    r = sw.SweepResult("learned", recs, 1.2)
    t = r.totals()
    assert t["good_pairs"] == 14 and t["factorizations"] == 3 and t["fallbacks"] == 1
    assert abs(r.eigenpairs_per_s - 14 / 1.2) < 1e-12
