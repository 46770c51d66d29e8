"""GPU parity of the device-resident advection + remap step (fv3_step_dev).

FUSED64 and PLAIN64 must match the FP64 oracle (tracer_2d_1L followed by the
tracer/pt remap) to 1e-12 relative per interior cell; FUSED32 must stay
within 1e-5 relative L2 per tracer and conserve tracer mass to 1e-6
relative over the advection.
"""
import numpy as np
import pytest
import torch

import fv3atm_amd as fa
from fv3atm_amd import cases
from oracle import oracle
from tests._util import assert_close, interior, rel_l2

pytestmark = pytest.mark.gpu
NG = 3


def _dev(a, dtype):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a).ravel(order="F"))).to(
        device="cuda", dtype=dtype)


def run_step(case, form, nsplit_cap=0, rows=0, cfg=0, concurrent=True, per_level=True, q_split=0,
             pipeline=None):
    dt = torch.float32 if form == fa.FORM_FUSED32 else torch.float64
    ctx = fa.Context(case.grids, case.npz)
    if rows or cfg:
        fa.set_launch(ctx, cfg, rows, 0)
    fa.set_concurrency(ctx, concurrent)
    if pipeline is not None:
        fa.set_pipeline(ctx, pipeline)
    fa.set_transport(ctx, per_level, q_split)
    q_in = _dev(case.q, dt)
    q_out = torch.zeros_like(q_in)
    pt = _dev(case.pt, dt) if case.pt is not None else None
    dp1 = _dev(case.dp1, torch.float64)
    delp_out = torch.zeros_like(dp1)
    cx, cy, mfx, mfy = (_dev(getattr(case, k), dt) for k in ("cx", "cy", "mfx", "mfy"))
    ak, bk = _dev(case.ak, torch.float64), _dev(case.bk, torch.float64)
    fa.step_dev(ctx, form, case.nq, q_in, q_out, pt, dp1, delp_out, cx, cy, mfx, mfy, ak, bk,
                case.ptop, case.hord, case.kord_tr, case.kord_tm, 184.0, nsplit_cap)
    torch.cuda.synchronize()
    shape_q = case.q.shape
    out = dict(q=q_out.double().cpu().numpy().reshape(shape_q, order="F"),
               delp=delp_out.cpu().numpy().reshape(case.dp1.shape, order="F"))
    if pt is not None:
        out["pt"] = pt.double().cpu().numpy().reshape(case.pt.shape, order="F")
    ctx.close()
    return out


def reference(case):
    adv = oracle.tracer_2d(case, "1L")
    q, delp, pt = oracle.remap_tracers(case, adv["q"], case.dp1, case.pt)
    return dict(q=q, delp=delp, pt=pt, q_adv=adv["q"], nsplit=adv["nsplit"])


CASES = {
    "C48L32Q1": lambda: cases.config(0, with_pt=True),
    "C24L12Q3jet": lambda: cases.make_case("C24L12Q3jet", n=24, npz=12, nq=3, jet=True,
                                           courant=2.5, delp_kind=1, delp_amp=(0.05, 0.2),
                                           seed=3, with_pt=True),
    # two strips, two row blocks, 7 tracers (mapn_tracer branch), uneven nsplit
    "C100L6Q7jet": lambda: cases.make_case("C100L6Q7", n=100, npz=6, nq=7, jet=True,
                                           courant=1.9, delp_kind=1, delp_amp=(0.05, 0.15),
                                           seed=11, with_pt=True),
    # the cubed sphere: edge halos and corners read from the neighbour faces
    "6xC32L8Q2": lambda: cases.make_case("6xC32L8Q2", n=32, npz=8, nq=2, ntiles=6, courant=1.7,
                                         delp_kind=1, delp_amp=(0.05, 0.15), seed=17, with_pt=True),
}


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("form", [fa.FORM_FUSED64, fa.FORM_PLAIN64])
def test_step_fp64_parity(gpu_ready, name, form):
    case = CASES[name]()
    ref = reference(case)
    print(name, "form", form, "nsplit", np.unique(ref["nsplit"]).tolist(), flush=True)
    got = run_step(case, form)
    assert_close(interior(got["q"]), interior(ref["q"]), 1e-12, f"{name} q", floor=1e-300)
    assert_close(interior(got["delp"]), interior(ref["delp"]), 1e-12, f"{name} delp")
    assert_close(interior(got["pt"]), interior(ref["pt"]), 1e-12, f"{name} pt")


@pytest.mark.parametrize("hord", [10, 13])
def test_step_fused64_hord(gpu_ready, hord):
    case = CASES["C24L12Q3jet"]()
    case.hord = hord
    ref = reference(case)
    got = run_step(case, fa.FORM_FUSED64)
    assert_close(interior(got["q"]), interior(ref["q"]), 1e-12, f"hord{hord} q", floor=1e-300)


@pytest.mark.parametrize("name", list(CASES))
def test_step_fused32_error(gpu_ready, name):
    case = CASES[name]()
    ref = reference(case)
    got = run_step(case, fa.FORM_FUSED32)
    errs = []
    for iq in range(case.nq):
        e = rel_l2(interior(got["q"][:, :, :, iq]), interior(ref["q"][:, :, :, iq]))
        errs.append(e)
    print(name, "FP32 rel L2 per tracer", ["%.2e" % e for e in errs], flush=True)
    assert max(errs) <= 1e-5
    assert rel_l2(interior(got["pt"]), interior(ref["pt"])) <= 1e-6
    assert_close(interior(got["delp"]), interior(ref["delp"]), 1e-12, "delp")


def test_step_nsplit_cap(gpu_ready):
    case = CASES["C24L12Q3jet"]()
    ref = reference(case)
    cap = int(ref["nsplit"].max())
    got = run_step(case, fa.FORM_FUSED64, nsplit_cap=cap + 1)
    assert_close(interior(got["q"]), interior(ref["q"]), 1e-12, "cap q", floor=1e-300)
    with pytest.raises(RuntimeError, match="nsplit_cap"):
        run_step(case, fa.FORM_FUSED64, nsplit_cap=cap - 1)


def test_step_interior_blocks(gpu_ready):
    # C192 L4 with 30-row blocks: strips 1-2 x blocks 1-5 of every level run the
    # lean interior kernel, the frame runs the cube-edge kernel; both must
    # reproduce the oracle bit for bit (FP64) and to 1e-5 (FP32)
    case = cases.make_case("C192L4Q2jet", n=192, npz=4, nq=2, jet=True, courant=2.1, delp_kind=1,
                           delp_amp=(0.05, 0.2), seed=13, with_pt=True)
    ref = reference(case)
    print("C192L4Q2jet nsplit", ref["nsplit"].tolist(), flush=True)
    got = run_step(case, fa.FORM_FUSED64, rows=30)
    assert_close(interior(got["q"]), interior(ref["q"]), 1e-12, "interior-block q", floor=1e-300)
    got32 = run_step(case, fa.FORM_FUSED32, rows=30)
    for iq in range(case.nq):
        e = rel_l2(interior(got32["q"][..., iq]), interior(ref["q"][..., iq]))
        print(f"FP32 tracer {iq}: rel L2 {e:.2e}")
        assert e <= 1e-5



def test_step_six_tiles_interior_blocks(gpu_ready):
    # six C128 faces with 30-row blocks: lean interior blocks on every face,
    # cube-edge frames reading their halos from the neighbour faces
    case = cases.make_case("6xC128L4Q2", n=128, npz=4, nq=2, ntiles=6, courant=2.2, delp_kind=1,
                           delp_amp=(0.05, 0.2), seed=19, with_pt=True)
    ref = reference(case)
    print("6xC128L4Q2 nsplit", ref["nsplit"].tolist(), flush=True)
    got = run_step(case, fa.FORM_FUSED64, rows=30)
    assert_close(interior(got["q"]), interior(ref["q"]), 1e-12, "six-tile q", floor=1e-300)
    got32 = run_step(case, fa.FORM_FUSED32, rows=30)
    for iq in range(case.nq):
        e = rel_l2(interior(got32["q"][..., iq, :]), interior(ref["q"][..., iq, :]))
        print(f"FP32 tracer {iq}: rel L2 {e:.2e}")
        assert e <= 1e-5


def test_step_every_launch_shape(gpu_ready):
    # every launch shape (waves per workgroup x tracers per wave, lean and
    # edge kinds, occupancy) on a case with all three block kinds and a
    # partial last round of tracers
    case = cases.make_case("C192L4Q7jet", n=192, npz=4, nq=7, jet=True, courant=2.1, delp_kind=1,
                           delp_amp=(0.05, 0.2), seed=23, with_pt=True)
    ref = reference(case)
    base = None
    for cfg in range(fa.N_LAUNCH_CFG):
        got = run_step(case, fa.FORM_FUSED32, rows=30, cfg=cfg)
        errs = [rel_l2(interior(got["q"][:, :, :, iq]), interior(ref["q"][:, :, :, iq])) for iq in range(case.nq)]
        print(f"cfg {cfg}: max rel L2 vs oracle {max(errs):.2e}", flush=True)
        assert max(errs) <= 1e-5
        if base is None:
            base = got["q"]
        else:
            assert rel_l2(interior(got["q"]), interior(base)) <= 1e-6


def test_step_two_rounds_default_shape(gpu_ready):
    # the context's default launch (16 tracers per lean round) over 19 tracers:
    # a full round, then one with a full wave, a wave with one inactive tracer
    # and idle waves; 50-row blocks leave a short last block (padded rows)
    case = cases.make_case("C96L4Q19jet", n=96, npz=4, nq=19, jet=True, courant=2.1, delp_kind=1,
                           delp_amp=(0.05, 0.2), seed=31, with_pt=True)
    print("[fused] C96 L4 Q19, default launch, 50 rows per block", flush=True)
    ref = reference(case)
    got = run_step(case, fa.FORM_FUSED32, rows=50, cfg=fa.DEFAULT_LAUNCH_CFG)
    errs = [rel_l2(interior(got["q"][:, :, :, iq]), interior(ref["q"][:, :, :, iq])) for iq in range(case.nq)]
    print(f"max rel L2 vs oracle {max(errs):.2e}", flush=True)
    assert max(errs) <= 1e-5
    assert rel_l2(interior(got["pt"]), interior(ref["pt"])) <= 1e-6


@pytest.mark.parametrize("form", [fa.FORM_FUSED32, fa.FORM_FUSED64])
def test_step_concurrent_kinds_equal_serial(gpu_ready, form):
    # the edge kernels on their own streams give the same bits as run after
    # the lean kernel on the step's stream (one tile and six tiles)
    for mk in (lambda: cases.make_case("C192L4Q3jet", n=192, npz=4, nq=3, jet=True, courant=2.1,
                                       delp_kind=1, delp_amp=(0.05, 0.2), seed=29, with_pt=True),
               lambda: cases.make_case("6xC128L4Q2", n=128, npz=4, nq=2, ntiles=6, courant=2.2,
                                       delp_kind=1, delp_amp=(0.05, 0.2), seed=31, with_pt=True)):
        case = mk()
        a = run_step(case, form, rows=30, concurrent=True)
        b = run_step(case, form, rows=30, concurrent=False)
        print(case.name, "max |concurrent - serial|", float(np.abs(a["q"] - b["q"]).max()), flush=True)
        assert np.array_equal(interior(a["q"]), interior(b["q"]))
        assert np.array_equal(interior(a["pt"]), interior(b["pt"]))


@pytest.mark.parametrize("chunk", [1, 3, -1])
def test_step_pipelined_chunks_equal_whole(gpu_ready, chunk):
    # tracers transported in chunks, each chunk remapped beside the next
    # chunk's transport: the same bits as all tracers at once (one tile and
    # six tiles; pt goes with the last chunk)
    for mk in (lambda: cases.make_case("C100L12Q7jet", n=100, npz=12, nq=7, jet=True, courant=1.9,
                                       delp_kind=1, delp_amp=(0.05, 0.15), seed=37, with_pt=True),
               lambda: cases.make_case("6xC48L8Q5", n=48, npz=8, nq=5, ntiles=6, courant=1.7,
                                       delp_kind=1, delp_amp=(0.05, 0.15), seed=41, with_pt=True)):
        case = mk()
        a = run_step(case, fa.FORM_FUSED32, pipeline=chunk, cfg=2 if chunk == -1 else 0)
        b = run_step(case, fa.FORM_FUSED32, pipeline=0, cfg=2 if chunk == -1 else 0)
        print(case.name, "chunk", chunk, "max |pipelined - whole|", float(np.abs(a["q"] - b["q"]).max()),
              flush=True)
        assert np.array_equal(interior(a["q"]), interior(b["q"]))
        assert np.array_equal(interior(a["pt"]), interior(b["pt"]))
        assert np.array_equal(interior(a["delp"]), interior(b["delp"]))


@pytest.mark.parametrize("q_split", [0, 3])
@pytest.mark.parametrize("name", ["C100L6Q7jet", "6xC32L8Q2"])
def test_step_tracer_2d_rule(gpu_ready, name, q_split):
    # tracer_2d sub-cycling (one nsplit for every level; forced to q_split when
    # given): FP64 to the oracle's tracer_2d bit for bit, FP32 to 1e-5
    case = CASES[name]()
    case.q_split = q_split
    adv = oracle.tracer_2d(case, "2d")
    q, delp, pt = oracle.remap_tracers(case, adv["q"], case.dp1, case.pt)
    print(name, "q_split", q_split, "nsplit", np.unique(adv["nsplit"]).tolist(), flush=True)
    for form in (fa.FORM_FUSED64, fa.FORM_PLAIN64):
        got = run_step(case, form, per_level=False, q_split=q_split)
        assert_close(interior(got["q"]), interior(q), 1e-12, f"{name} form {form} tracer_2d q", floor=1e-300)
        assert_close(interior(got["pt"]), interior(pt), 1e-12, f"{name} form {form} tracer_2d pt")
    got32 = run_step(case, fa.FORM_FUSED32, per_level=False, q_split=q_split)
    for iq in range(case.nq):
        e = rel_l2(interior(got32["q"][:, :, :, iq]), interior(q[:, :, :, iq]))
        assert e <= 1e-5, (iq, e)
