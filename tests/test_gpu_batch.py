"""GPU tests of the batched uneven launch, graph replay and autotuner (N7).

Per-instance results of one batched execution must equal the plain per-plan
results (rel-L2 <= 1e-6: the same kernels on the same work items), graph replays must equal eager execution bit for bit
(with the inputs rewritten in place and re-pointed between replays), and the
autotuned configuration must stay within 1e-5 of the FP64 oracle.
"""
import numpy as np
import pytest

from oracle import nufft_oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = a.detach().cpu().numpy() if hasattr(a, "detach") else np.asarray(a)
    b = b.detach().cpu().numpy() if hasattr(b, "detach") else np.asarray(b)
    a = a.astype(np.complex128).ravel()
    b = b.astype(np.complex128).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def small_batch(seed=7, n=12, dcf=True):
    from mripy_amd import data
    inst = data.config_batched_uneven(n, seed=seed, sizes=(24, 32, 40, 48), spokes=(8, 40), coils=(1, 12))
    rng = np.random.default_rng(seed)
    for k, ins in enumerate(inst):
        if dcf and k % 3 == 0:
            ins["dcf"] = rng.random(ins["coords"][0].shape[0]).astype(np.float32)
    return inst


def plain(inst, kind, src):
    from mripy_amd.plan import Plan
    P = Plan(inst["modes"], ncoils=inst["ncoils"], rat=2, nspread=3)
    P.set_points(inst["coords"], dcf=inst.get("dcf"))
    return getattr(P, kind)(src)


def test_batch_equals_plain_plans_all_kinds():
    import torch
    from mripy_amd.batch import BatchPlan
    print("case: batch vs plain, 12 uneven instances", flush=True)
    inst = small_batch()
    B = BatchPlan(inst, rat=2, nspread=3)
    imgs = [torch.as_tensor(i["image"]).cuda() for i in inst]
    rng = np.random.default_rng(1)
    ys = [torch.as_tensor((rng.standard_normal((i["coords"][0].shape[0], i["ncoils"])) +
                           1j * rng.standard_normal((i["coords"][0].shape[0], i["ncoils"]))).astype(np.complex64)).cuda()
          for i in inst]
    for kind, src in (("type2", imgs), ("type1", ys), ("adjoint", ys), ("type21", imgs)):
        got = getattr(B, kind)(src)
        for i, ins in enumerate(inst):
            assert rel(got[i], plain(ins, kind, src[i])) <= 1e-6, (kind, i)
    # two instances against the FP64 oracle
    for i in (0, 5):
        p = O.Params(inst[i]["modes"], rat=2, nspread=3)
        c64 = [c.astype(np.float64) for c in inst[i]["coords"]]
        assert rel(B.type2(imgs)[i], O.type2(c64, inst[i]["image"], p)) <= 1e-5
    info = B.info()
    # 4 grid lengths: one launch per length for the row pass and the column pass,
    # plus the gather launch shapes; 12 plain plans need 36
    assert info["instances"] == 12 and info["launches_type2"] <= 14


def test_graph_replays_equal_eager():
    import torch
    from mripy_amd.batch import BatchPlan
    from mripy_amd._lib import GN_TYPE1, GN_TYPE2
    print("case: graph replay vs eager", flush=True)
    inst = small_batch(seed=8, n=10)
    B = BatchPlan(inst, rat=2, nspread=3)
    G = BatchPlan(inst, rat=2, nspread=3, graph=True)
    rng = np.random.default_rng(2)
    imgs = [torch.as_tensor(i["image"]).cuda() for i in inst]
    ys = [torch.empty(B.sample_shape(i), dtype=torch.complex64, device="cuda") for i in range(B.n)]
    ims = [torch.empty(B.image_shape(i), dtype=torch.complex64, device="cuda") for i in range(B.n)]
    ys_g = [torch.empty_like(t) for t in ys]
    ims_g = [torch.empty_like(t) for t in ims]
    for rep in range(4):
        if rep == 2:   # new input pointers
            imgs = [t.clone() for t in imgs]
        for t in imgs:   # new content at the same pointers
            t.copy_(torch.as_tensor((rng.standard_normal(t.shape) + 1j * rng.standard_normal(t.shape))
                                    .astype(np.complex64)))
        B.execute_into(GN_TYPE2, imgs, ys)
        B.execute_into(GN_TYPE1, ys, ims)
        G.execute_into(GN_TYPE2, imgs, ys_g)
        G.execute_into(GN_TYPE1, ys_g, ims_g)
        torch.cuda.synchronize()
        _ = (imgs[0] * 2).sum().item()   # eager work between replays
        for i in range(B.n):
            assert torch.equal(ys[i], ys_g[i]), (rep, i)
            assert torch.equal(ims[i], ims_g[i]), (rep, i)


def test_batch_autotune_and_bad_tiles():
    import torch
    from mripy_amd.batch import BatchPlan
    from mripy_amd._lib import GN_TYPE1
    print("case: batch autotune", flush=True)
    inst = small_batch(seed=9, n=6, dcf=False)
    B = BatchPlan(inst, rat=2, nspread=3)
    rng = np.random.default_rng(3)
    ys = [torch.as_tensor((rng.standard_normal(B.sample_shape(i)) + 1j * rng.standard_normal(B.sample_shape(i)))
                          .astype(np.complex64)).cuda() for i in range(B.n)]
    ims = [torch.empty(B.image_shape(i), dtype=torch.complex64, device="cuda") for i in range(B.n)]
    (sp, ga), ms = B.autotune(GN_TYPE1, ys, ims, reps=2)
    assert ms > 0 and B.info()["tiles"] == sp and B.info()["gather_tiles"] == ga
    out = B.type1(ys)
    for i in (0, 3):
        p = O.Params(inst[i]["modes"], rat=2, nspread=3)
        c64 = [c.astype(np.float64) for c in inst[i]["coords"]]
        assert rel(out[i], O.type1(c64, ys[i].cpu().numpy(), p)) <= 1e-5
    # a gather tile whose staged region exceeds the LDS budget is refused, and the
    # batch keeps its previous shape and results
    from mripy_amd import data
    big = small_batch(seed=9, n=3, dcf=False) + data.config_batched_uneven(1, seed=11, sizes=(128,), spokes=(16, 16),
                                                                             coils=(2, 2))
    C = BatchPlan(big, rat=2, nspread=3)
    ys2 = [torch.as_tensor((rng.standard_normal(C.sample_shape(i)) + 1j * rng.standard_normal(C.sample_shape(i)))
                           .astype(np.complex64)).cuda() for i in range(C.n)]
    before = C.type1(ys2)
    shape = C.info()
    with pytest.raises(ValueError):
        C.set_tiles((16, 32), (200, 200))       # (200+5)^2 complex64 > 160 KiB of LDS
    assert C.info()["tiles"] == shape["tiles"] and C.info()["gather_tiles"] == shape["gather_tiles"]
    after = C.type1(ys2)
    for i in range(C.n):
        assert torch.equal(before[i], after[i])


def test_plan_autotune_3d():
    import torch
    from mripy_amd.plan import Plan
    from mripy_amd._lib import GN_TYPE2
    print("case: plan autotune 3-D", flush=True)
    rng = np.random.default_rng(4)
    M, modes = 20000, (20, 16, 24)
    cs = [(np.pi * (2 * rng.random(M) - 1)).astype(np.float32) for _ in range(3)]
    P = Plan(modes, ncoils=2, rat=2, nspread=3).set_points(cs)
    img = torch.as_tensor((rng.standard_normal(modes + (2,)) + 1j * rng.standard_normal(modes + (2,)))
                          .astype(np.complex64)).cuda()
    y = torch.empty((M, 2), dtype=torch.complex64, device="cuda")
    (sp, ga), ms = P.autotune(GN_TYPE2, img, y, reps=1)
    assert ms > 0 and P.tiles == sp and P.gather_tiles == ga
    p = O.Params(modes, rat=2, nspread=3)
    assert rel(P.type2(img), O.type2([c.astype(np.float64) for c in cs], img.cpu().numpy(), p)) <= 1e-5


def test_batch_rejects_type1_of_empty_instance():
    from mripy_amd.batch import BatchPlan
    inst = small_batch(seed=10, n=3, dcf=False)
    inst[1]["coords"] = [np.zeros(0, np.float32), np.zeros(0, np.float32)]
    B = BatchPlan(inst, rat=2, nspread=3)
    with pytest.raises(ValueError):
        B.type1([np.zeros(B.sample_shape(i), np.complex64) for i in range(3)])
    out = B.type2([i["image"] for i in inst])
    assert out[1].shape == (0, inst[1]["ncoils"])


def test_uneven_64_instance_workload_batched_equals_plain():
    import torch
    from mripy_amd import data
    from mripy_amd.batch import BatchPlan
    print("case: 64 uneven instances, N in {128..320}, nc in 1..32", flush=True)
    inst = data.config_batched_uneven()
    B = BatchPlan(inst, rat=2, nspread=3, graph=True)
    imgs = [torch.as_tensor(i["image"]).cuda() for i in inst]
    ys = B.type2(imgs)
    ims = B.type1(ys)
    worst2 = worst1 = 0.0
    for i, ins in enumerate(inst):
        y_p = plain(ins, "type2", imgs[i])
        worst2 = max(worst2, rel(ys[i], y_p))
        worst1 = max(worst1, rel(ims[i], plain(ins, "type1", ys[i])))
    print("max rel vs plain: type2 %.3g type1 %.3g" % (worst2, worst1), flush=True)
    assert worst2 <= 1e-5 and worst1 <= 1e-5


def test_batch_3d_equals_plain():
    """3-D instances of three grid sizes: the batched pruned FFT runs passes A, B, C
    (and C', B', A') over all instances; every kind equals the plain plans."""
    import torch
    from mripy_amd.batch import BatchPlan
    print("case: 3-D batch vs plain", flush=True)
    rng = np.random.default_rng(21)
    inst = []
    for modes, nc in (((16, 16, 16), 3), ((24, 20, 16), 1), ((16, 16, 16), 9), ((20, 24, 12), 5)):
        M = 3000 + 500 * nc
        coords = [(np.pi * (2 * rng.random(M) - 1)).astype(np.float32) for _ in range(3)]
        img = (rng.standard_normal(modes + (nc,)) + 1j * rng.standard_normal(modes + (nc,))).astype(np.complex64)
        inst.append({"modes": modes, "coords": coords, "ncoils": nc, "image": img})
    B = BatchPlan(inst, rat=2, nspread=3)
    imgs = [torch.as_tensor(i["image"]).cuda() for i in inst]
    ys = [torch.as_tensor((rng.standard_normal((i["coords"][0].shape[0], i["ncoils"])) +
                           1j * rng.standard_normal((i["coords"][0].shape[0], i["ncoils"]))).astype(np.complex64)).cuda()
          for i in inst]
    for kind, src in (("type2", imgs), ("type1", ys), ("type21", imgs)):
        got = getattr(B, kind)(src)
        for i, ins in enumerate(inst):
            assert rel(got[i], plain(ins, kind, src[i])) <= 1e-6, (kind, i)
    p = O.Params(inst[1]["modes"], rat=2, nspread=3)
    c64 = [c.astype(np.float64) for c in inst[1]["coords"]]
    assert rel(B.type2(imgs)[1], O.type2(c64, inst[1]["image"], p)) <= 1e-5


def test_batch_mixes_matrix_core_and_vector_spreads():
    """A tiny instance (grid 12 x 16: its spread tiles shrink to fit and are not
    block-aligned) runs the vector spread inside the same batch as matrix-core ones."""
    import torch
    from mripy_amd.batch import BatchPlan
    print("case: batch with a vector-spread instance", flush=True)
    rng = np.random.default_rng(23)
    inst = []
    for modes, nc in (((6, 8), 2), ((32, 40), 3), ((6, 8), 9), ((24, 24), 1)):
        M = 400 + 100 * nc
        coords = [(np.pi * (2 * rng.random(M) - 1)).astype(np.float32) for _ in range(2)]
        inst.append({"modes": modes, "coords": coords, "ncoils": nc})
    B = BatchPlan(inst, rat=2, nspread=3)
    ys = [torch.as_tensor((rng.standard_normal((i["coords"][0].shape[0], i["ncoils"])) +
                           1j * rng.standard_normal((i["coords"][0].shape[0], i["ncoils"]))).astype(np.complex64)).cuda()
          for i in inst]
    got = B.type1(ys)
    for i, ins in enumerate(inst):
        assert rel(got[i], plain(ins, "type1", ys[i])) <= 1e-6, i
    p = O.Params(inst[0]["modes"], rat=2, nspread=3)
    c64 = [c.astype(np.float64) for c in inst[0]["coords"]]
    assert rel(got[0], O.type1(c64, ys[0].cpu().numpy(), p)) <= 1e-5
