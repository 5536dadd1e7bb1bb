"""Graph replay of steady substeps, launch schedules, the autotuner and concurrent batches of
uneven size classes (crm_hip_run_many).  None of these may change a result: every check is
bitwise equality with the plain eager call of the same class, plus one parity check of a
batched class against the oracle."""
import numpy as np
import pytest

import oracle_lib
from crm_hip import KERNELS, CrmHip, DeviceCase, run_many, tune_many
from crm_inputs import make_case
from parity import compare_case, convective_case

pytestmark = pytest.mark.gpu

CLASSES = [  # (ncrms, nx, ny, nz, dx, dt): the uneven-batch size classes, scaled down
    (64, 16, 1, 28, 4000.0, 20.0),
    (32, 32, 1, 28, 4000.0, 20.0),
    (16, 64, 1, 58, 1000.0, 5.0),
    (8, 8, 8, 28, 4000.0, 20.0),
    (3, 128, 1, 58, 1000.0, 5.0),
]


def _case(cls, seed):
    n, nx, ny, nz, dx, dt = cls
    return convective_case(n, nx=nx, ny=ny, nz=nz, dx=dx, dt=dt, seed=seed)


def _same(a, b):
    for g in ("state", "rad", "out"):
        for k, v in getattr(a, g).items():
            assert np.array_equal(getattr(b, g)[k], v), (g, k)


@pytest.mark.parametrize("cls", CLASSES, ids=lambda c: f"{c[0]}x{c[1]}x{c[2]}x{c[3]}")
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("mode", [1, 2], ids=["substep", "call"])
def test_graphs_replay_bitwise(cls, prec, mode):
    """Three consecutive calls (the graphs of the first are replayed by the later ones, and
    each call replays its graphs many times) against the same calls run eagerly."""
    case = _case(cls, seed=81)
    eager, graphed = case.copy(), case.copy()
    he, hg = CrmHip(0, prec), CrmHip(0, prec)
    try:
        he.set_graphs(False)
        hg.set_graphs(mode)
        for _ in range(3):
            he.run(eager, 10 * cls[5])
            hg.run(graphed, 10 * cls[5])
            _same(eager, graphed)
        assert he.graph_launches() == 0
        assert hg.graph_launches() > 0
    finally:
        he.close()
        hg.close()


def test_schedule_does_not_change_results():
    case = _case(CLASSES[0], seed=82)
    a, b = case.copy(), case.copy()
    h = CrmHip(0)
    try:
        h.run(a, 100.0)
        h.set_schedule({k: 64 for k in KERNELS})
        assert set(h.schedule().values()) == {64}
        h.run(b, 100.0)
        _same(a, b)
    finally:
        h.close()


def test_autotune_schedule():
    """The tuner returns one of its candidate sizes for every class and leaves a context whose
    results are those of the default schedule."""
    cls = CLASSES[1]
    case = _case(cls, seed=83)
    h, ref = CrmHip(0), CrmHip(0)
    try:
        scratch = DeviceCase(case.copy())
        sched = h.autotune(scratch, 3 * cls[5])
        print(sched)
        assert set(sched.values()) <= {64, 128, 256}
        assert h.schedule() == sched
        a, b = case.copy(), case.copy()
        h.run(a, 100.0)
        ref.run(b, 100.0)
        _same(a, b)
    finally:
        h.close()
        ref.close()


@pytest.mark.parametrize("cls", [CLASSES[2], (8, 64, 1, 28, 4000.0, 20.0), CLASSES[1], CLASSES[3]],
                         ids=lambda c: f"{c[0]}x{c[1]}x{c[2]}x{c[3]}")
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_launch_variants_do_not_change_results(cls, prec):
    """Every launch variant of the pressure solve (4-wave reading its inputs twice, one, two or
    four levels per trip, Thomas factors in LDS or in registers; 8-wave with the inputs in
    registers) and of MPDATA (one or two levels per trip, registers bounded for two or three
    waves per SIMD, the scalars in one sweep or on separate waves) gives the default variant's
    bits."""
    case = _case(cls, seed=87)
    ref = case.copy()
    h = CrmHip(0, prec)
    try:
        h.run(ref, 4 * cls[5])
        nv = 5 if (cls[1] == 64 and cls[2] == 1) else 1
        for press in range(nv):
            for mp in range(6):
                h.set_variants({"press": press, "mpdata": mp})
                assert h.variants()["press"] == press and h.variants()["mpdata"] == mp
                got = case.copy()
                h.run(got, 4 * cls[5])
                _same(ref, got)
    finally:
        h.close()


def test_autotune_chooses_variants():
    """At 64 columns the tuner measures every pressure and MPDATA variant and keeps one of
    them; results stay bitwise those of the default."""
    cls = CLASSES[2]
    case = _case(cls, seed=88)
    h, ref = CrmHip(0), CrmHip(0)
    try:
        h.autotune(DeviceCase(case.copy()), 3 * cls[5])
        v = h.variants()
        print(v)
        assert v["press"] in (0, 1, 2, 3, 4) and v["mpdata"] in (0, 1, 2, 3, 4, 5)
        assert all(x == 0 for k, x in v.items() if k not in ("press", "mpdata"))
        a, b = case.copy(), case.copy()
        h.run(a, 100.0)
        ref.run(b, 100.0)
        _same(a, b)
    finally:
        h.close()
        ref.close()


def test_run_many_matches_separate_calls():
    """Four uneven classes advanced concurrently, each equal to its own separate call."""
    cases = [_case(c, seed=84 + j) for j, c in enumerate(CLASSES)]
    alone = []
    for c in cases:
        h = CrmHip(0)
        try:
            a = c.copy()
            h.run(a, 200.0)
            alone.append(a)
        finally:
            h.close()
    hips = [CrmHip(0) for _ in CLASSES]
    try:
        dcs = [DeviceCase(c.copy()) for c in cases]
        status = run_many(hips, dcs, 200.0)
        assert status == [0] * len(CLASSES)
        for dc, a in zip(dcs, alone):
            _same(dc.download(dc.case), a)
    finally:
        for h in hips:
            h.close()


def test_tune_many_restores_state_and_keeps_results():
    """The batch tuner tries every pressure and MPDATA variant each class has, leaves the
    DeviceCases in the state they had on entry, and the batch it tuned stays bitwise equal to
    separate calls with the default variants."""
    cls = [CLASSES[2], CLASSES[4]]  # 64 columns (5 pressure variants) and 128 columns (1)
    cases = [_case(c, seed=90 + j) for j, c in enumerate(cls)]
    alone = []
    for c in cases:
        h = CrmHip(0)
        try:
            a = c.copy()
            h.run(a, 200.0)
            alone.append(a)
        finally:
            h.close()
    hips = [CrmHip(0) for _ in cls]
    try:
        dcs = [DeviceCase(c.copy()) for c in cases]
        chosen, trials = tune_many(hips, dcs, 200.0, kinds=("press", "mpdata"))
        assert trials == 1 + 4 + 5 + 5
        assert chosen[0]["press"] in range(5) and chosen[1]["press"] == 0
        assert all(v["mpdata"] in range(6) for v in chosen)
        for dc, c in zip(dcs, cases):
            _same(c, dc.download(c.copy()))
        assert run_many(hips, dcs, 200.0) == [0, 0]
        for dc, a in zip(dcs, alone):
            _same(dc.download(dc.case), a)
    finally:
        for h in hips:
            h.close()


@pytest.mark.parametrize("mode", [2, 3], ids=["per_class_graphs", "joint_graph"])
def test_run_many_one_graph(mode):
    """crm_hip_run_many with CRM_GRAPHS_JOINT captures the classes' whole calls into one graph
    (cached, so the second and third calls replay it); with CRM_GRAPHS_CALL each class replays
    its own graph on its own stream.  Quiet soundings run every class's graph to the end, a
    class stopped mid-call by the test hook resumes on its own, a class with mean-state
    acceleration runs outside the graph, and a mixed-precision class joins; every class stays
    bitwise equal to its own separate call, call after call."""
    cls = [CLASSES[0], CLASSES[1], CLASSES[2], CLASSES[3], CLASSES[1]]
    precs = ["fp64", "fp64", "mixed", "fp64", "fp64"]
    cases = [make_case(c[0], c[1], c[3], c[4], c[5], ny=c[2], randomize=True, seed_base=200 + j)
             for j, c in enumerate(cls)]

    def configure(h, j):
        if j == 1:
            h.debug_stop_call_at(3)
        if j == 4:
            h.accel_init(True, 1.0)

    alone = []
    for j, c in enumerate(cases):
        h = CrmHip(0, precs[j])
        try:
            configure(h, j)
            a = c.copy()
            outs = []
            for _ in range(3):
                h.run(a, 200.0)
                outs.append(a.copy())
            alone.append(outs)
        finally:
            h.close()
    hips = [CrmHip(0, p) for p in precs]
    try:
        for j, h in enumerate(hips):
            h.set_graphs(mode)
            configure(h, j)
        dcs = [DeviceCase(c.copy()) for c in cases]
        for step in range(3):
            assert run_many(hips, dcs, 200.0) == [0] * len(cls)
            for dc, a in zip(dcs, alone):
                _same(dc.download(dc.case.copy()), a[step])
        assert hips[0].graph_launches() == 3 * 10  # quiet: every substep from the graph
        assert hips[4].graph_launches() == 0       # acceleration: plain launches
    finally:
        for h in hips:
            h.close()


def test_run_many_parity():
    """Batched 2D classes against the oracle (FP64 bars: 1e-10 after one substep, 1e-8 after
    several)."""
    cls = [CLASSES[0], CLASSES[2]]
    cases = [make_case(c[0], c[1], c[3], c[4], c[5], randomize=True, seed_base=90) for c in cls]
    hips = [CrmHip(0) for _ in cls]
    try:
        dcs = [DeviceCase(c.copy()) for c in cases]
        assert run_many(hips, dcs, 20.0) == [0, 0]
        for dc, c in zip(dcs, cases):
            want = c.copy()
            assert oracle_lib.oracle().run(want, 20.0) == 0
            errs = compare_case(dc.download(dc.case), want, dt_gl=20.0)
            tol = 1e-10 if c.dt == 20.0 else 1e-8  # one substep, or four
            assert max(errs.values()) <= tol, sorted(errs.items(), key=lambda kv: -kv[1])[:3]
    finally:
        for h in hips:
            h.close()



@pytest.mark.parametrize("stop_at", [1, 2, 3, 7, 10])
def test_call_graph_resumes_mid_call(stop_at):
    """A captured call stopped at any substep (the stop of a substep that needs more subcycles,
    forced here by the test hook) resumes on the host from its record of the state before that
    substep: the result is bitwise the eager call's."""
    case = _case(CLASSES[1], seed=85)
    eager, graphed = case.copy(), case.copy()
    he, hg = CrmHip(0), CrmHip(0)
    try:
        he.set_graphs(False)
        hg.debug_stop_call_at(stop_at)
        for _ in range(2):
            he.run(eager, 10 * CLASSES[1][5])
            hg.run(graphed, 10 * CLASSES[1][5])
            _same(eager, graphed)
    finally:
        he.close()
        hg.close()


def test_graph_cache_is_bounded():
    """Every batch size a context sees gets its own captured call graph (the batch size is one
    of the kernel arguments); the context keeps at most 32 of them, the least recently used
    evicted, and results stay bitwise those of plain launches."""
    n0 = 40
    # quiet soundings: one subcycle per substep, so each call replays its graph to the end
    base = make_case(n0, 32, 28, 4000.0, 20.0, randomize=True, seed_base=86)
    hg, he = CrmHip(0), CrmHip(0)
    try:
        he.set_graphs(False)
        for j, n in enumerate(range(n0, 0, -1)):  # largest first: the buffers are sized once
            case = base.copy()
            case.ncrms = n
            want = case.copy()
            hg.run(case, 40.0)
            assert hg.graph_cache_size() <= 32
            if j % 13 == 0 or n == 1:
                he.run(want, 40.0)
                _same(want, case)
        assert hg.graph_cache_size() == 32
        assert hg.graph_launches() >= 2 * n0
    finally:
        hg.close()
        he.close()
