"""H_eff matvec on the MI355X through the C ABI versus the oracle (GPU)."""
import sys

import numpy as np
import pytest

from oracle import heff as oheff
from su2dmrg_amd import synthetic

pytestmark = pytest.mark.gpu

RTOL = 1e-10   # relative 2-norm error bound for FP64 matvec parity


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    from su2dmrg_amd import lib as L
    c = L.Context(0)
    yield c
    c.close()


def _relerr(got, ref):
    num = sum(float(np.sum((got[j] - ref[j]) ** 2)) for j in ref)
    den = sum(float(np.sum(ref[j] ** 2)) for j in ref)
    return np.sqrt(num / den)


def _run(ctx, work):
    from su2dmrg_amd import lib as L
    dev = L.upload_workload(ctx, work)
    dev["plan"].apply(dev["psi"], dev["out"])
    got = L.unpack_wavefunction(dev["out"], dev["out"].download())
    return dev, got


@pytest.mark.parametrize("seed,twoj,dims,nterm", [
    (0, (0, 1, 2, 3, 4), (3, 4, 5, 2, 1), 3),
    (1, (1, 3, 5), (17, 33, 9), 2),
    (2, (0, 2, 4, 6), (130, 257, 129, 7), 4),
    (3, (1, 3, 5, 7), (300, 161, 64, 1), 1),
])
def test_heff_small_matches_oracle(ctx, seed, twoj, dims, nterm):
    _say(f"heff small seed={seed} twoj={twoj} dims={dims} nterm={nterm}")
    work = synthetic.small(seed=seed, twoj=twoj, dims=dims, nterm=nterm)
    _, got = _run(ctx, work)
    terms = [(L, R) for L, R in work["terms"]]
    ref = oheff.heff_apply(work["psi"], work["hl"], work["hr"], terms)
    assert _relerr(got, ref) < RTOL


def test_heff_cyl8x4_d1000_matches_oracle(ctx):
    _say("heff cyl8x4_d1000 seed=0")
    work = synthetic.cyl8x4_d1000(seed=0)
    dev, got = _run(ctx, work)
    ref = oheff.heff_apply(work["psi"], work["hl"], work["hr"], work["terms"])
    assert _relerr(got, ref) < RTOL
    alg, _ = dev["plan"].flops()
    assert alg == synthetic.flops(work)


def test_heff_without_block_hamiltonians_and_sparse_ops(ctx):
    _say("heff sparse ops, no hl/hr")
    work = synthetic.small(seed=11, twoj=(0, 2, 4, 6), dims=(20, 40, 30, 5), nterm=3)
    # drop some transitions from some terms: the plan zero-fills the bundle
    L0, R0 = work["terms"][0]
    del L0[(2, 4)]
    L1, R1 = work["terms"][1]
    del R1[(4, 2)]
    work["hl"] = None
    work["hr"] = None
    _, got = _run(ctx, work)
    ref = oheff.heff_apply(work["psi"], None, None, work["terms"])
    assert _relerr(got, ref) < RTOL


def test_heff_presummed_sides(ctx):
    from su2dmrg_amd import lib as L
    _say("heff presum")
    work = synthetic.small(seed=21, twoj=(1, 3, 5), dims=(40, 60, 20), nterm=3)
    twoj, da, db = work["twoj"], work["dims_a"], work["dims_b"]
    (L0, R0), (L1, R1), (L2, R2) = work["terms"]
    psi = L.wavefunction(ctx, twoj, da, db, work["psi"])
    out = L.wavefunction(ctx, twoj, da, db)
    tl = [L.vector_operator(ctx, twoj, da, x) for x in (L0, L1)]
    tr = [L.vector_operator(ctx, twoj, db, x) for x in (R0, R1, R2)]
    plan = L.HeffPlan(ctx, psi, None, None,
                      [([(1.0, tl[0])], [(0.5, tr[0]), (-2.0, tr[1])]),
                       ([(0.3, tl[0]), (1.7, tl[1])], [(1.0, tr[2])])])
    plan.apply(psi, out)
    got = L.unpack_wavefunction(out, out.download())
    terms = [(L0, oheff.presum([(0.5, R0), (-2.0, R1)])),
             (oheff.presum([(0.3, L0), (1.7, L1)]), R2)]
    ref = oheff.heff_apply(work["psi"], None, None, terms)
    assert _relerr(got, ref) < RTOL


def test_heff_many_summands(ctx):
    # Both terms share the interleaved right bundle of each transition, whose
    # summand count is the longer side (17): more than the bundle gather holds
    # in registers per term, so the bundle is assembled element by element,
    # with the shorter side's missing summands as absent entries.
    from su2dmrg_amd import lib as L
    _say("heff with 17 right summands")
    work = synthetic.small(seed=23, twoj=(1, 3, 5), dims=(40, 60, 20), nterm=18)
    twoj, da, db = work["twoj"], work["dims_a"], work["dims_b"]
    Ls = [t[0] for t in work["terms"]]
    Rs = [t[1] for t in work["terms"]]
    psi = L.wavefunction(ctx, twoj, da, db, work["psi"])
    out = L.wavefunction(ctx, twoj, da, db)
    tl = [L.vector_operator(ctx, twoj, da, x) for x in Ls[:2]]
    tr = [L.vector_operator(ctx, twoj, db, x) for x in Rs]
    rng = np.random.default_rng(7)
    c0 = list(rng.uniform(-1.5, 1.5, 17))
    c1 = list(rng.uniform(-1.5, 1.5, 3))
    plan = L.HeffPlan(ctx, psi, None, None,
                      [([(1.0, tl[0])], [(c, t) for c, t in zip(c0, tr[:17])]),
                       ([(1.0, tl[1])], [(c, t) for c, t in zip(c1, tr[15:18])])])
    plan.apply(psi, out)
    got = L.unpack_wavefunction(out, out.download())
    terms = [(Ls[0], oheff.presum(list(zip(c0, Rs[:17])))),
             (Ls[1], oheff.presum(list(zip(c1, Rs[15:18]))))]
    ref = oheff.heff_apply(work["psi"], None, None, terms)
    assert _relerr(got, ref) < RTOL


def test_heff_deterministic(ctx):
    from su2dmrg_amd import lib as L
    _say("heff determinism")
    work = synthetic.small(seed=5, twoj=(0, 2, 4), dims=(300, 500, 200), nterm=4)
    dev = L.upload_workload(ctx, work)
    dev["plan"].apply(dev["psi"], dev["out"])
    a = dev["out"].download()
    dev["plan"].apply(dev["psi"], dev["out"])
    b = dev["out"].download()
    assert np.array_equal(a, b)


def test_every_tile_variant_matches_oracle(ctx):
    from su2dmrg_amd import lib as L
    _say("heff tile variants")
    work = synthetic.small(seed=13, twoj=(1, 3, 5, 7), dims=(150, 290, 77, 3), nterm=3)
    dev = L.upload_workload(ctx, work)
    ref = oheff.heff_apply(work["psi"], work["hl"], work["hr"], work["terms"])
    chosen, t1, t2 = dev["plan"].autotune(reps=1)
    assert len(t1) == len(t2) >= 2 and all(t > 0 for t in t1 + t2)
    for v in range(len(t1)):
        _say(f"  variant {v}")
        dev["plan"].set_variants(v, v)
        dev["plan"].apply(dev["psi"], dev["out"])
        got = L.unpack_wavefunction(dev["out"], dev["out"].download())
        assert _relerr(got, ref) < RTOL, v


def test_uniform_and_general_gemm_kernels_agree_bitwise(ctx):
    # stage 1 is all op(A) = A^T, stage 2 all op(A) = A (H_L is transposed at
    # plan build), so both launches take the straight-line kernel; it stages
    # the same values through LDS in the same order as the general kernel,
    # so the two agree to the bit for every tile shape (M / N edges, K tails
    # and split K included by these dimensions)
    from su2dmrg_amd import lib as L
    _say("heff uniform vs general GEMM kernel")
    work = synthetic.small(seed=21, twoj=(1, 3, 5, 7), dims=(150, 290, 77, 3), nterm=3)
    dev = L.upload_workload(ctx, work)
    ref = oheff.heff_apply(work["psi"], work["hl"], work["hr"], work["terms"])
    plan = dev["plan"]
    _, t1, _ = plan.autotune(reps=1)
    for v in range(len(t1)):
        _say(f"  variant {v}")
        plan.set_variants(v, v)
        assert plan.set_kernel_path(True) == (1, 0)
        plan.apply(dev["psi"], dev["out"])
        a = dev["out"].download()
        assert plan.set_kernel_path(False) == (-1, -1)
        plan.apply(dev["psi"], dev["out"])
        b = dev["out"].download()
        assert np.array_equal(a, b), v
        got = L.unpack_wavefunction(dev["out"], a)
        assert _relerr(got, ref) < RTOL, v
    plan.set_kernel_path(True)


def test_vector_algebra(ctx):
    from su2dmrg_amd import lib as L
    _say("vector algebra")
    work = synthetic.small(seed=9, twoj=(1, 3), dims=(70, 33), nterm=0)
    twoj, da, db = work["twoj"], work["dims_a"], work["dims_b"]
    rng = np.random.default_rng(3)
    u = {j: rng.standard_normal(b.shape) for j, b in work["psi"].items()}
    v = {j: rng.standard_normal(b.shape) for j, b in work["psi"].items()}
    tu = L.wavefunction(ctx, twoj, da, db, u)
    tv = L.wavefunction(ctx, twoj, da, db, v)
    assert abs(tu.dot(tv) - oheff.inner(u, v)) < 1e-12 * oheff.norm(u) * oheff.norm(v)
    tu.axpy(-0.25, tv)
    ref = oheff.axpy(-0.25, v, u)
    got = L.unpack_wavefunction(tu, tu.download())
    assert _relerr(got, ref) < 1e-14
    n = tu.normalize()
    assert abs(n - oheff.norm(ref)) < 1e-12 * n
    tw = L.wavefunction(ctx, twoj, da, db)
    tw.copy_from(tu)
    tw.scal(3.0)
    assert abs(tw.dot(tw) - 9.0) < 1e-12


@pytest.fixture(scope="module")
def headline():
    """16x8 D=2000 centre-bond workload (the benchmarked matvec) and its oracle result."""
    work = synthetic.cyl16x8_d2000(seed=0)
    ref = oheff.heff_apply(work["psi"], work["hl"], work["hr"], work["terms"])
    return work, ref


def test_heff_cyl16x8_d2000_default_variants_match_oracle(ctx, headline):
    # reaches what only the headline size does: stage 2 split into ~4000 work
    # units, ~8000 workgroups through the XCD remap, 32-bit byte offsets on
    # 10896 x 1362 operands
    _say("heff cyl16x8_d2000 seed=0, default tile variants")
    work, ref = headline
    dev, got = _run(ctx, work)
    assert _relerr(got, ref) < RTOL
    alg, _ = dev["plan"].flops()
    assert alg == synthetic.flops(work)


def test_heff_cyl16x8_d2000_autotuned_matches_oracle(ctx, headline):
    from su2dmrg_amd import lib as L
    _say("heff cyl16x8_d2000 seed=0, autotuned tile variants")
    work, ref = headline
    dev = L.upload_workload(ctx, work)
    chosen, _, _ = dev["plan"].autotune(reps=1)
    _say(f"  autotuned variants {chosen}")
    dev["plan"].apply(dev["psi"], dev["out"])
    got = L.unpack_wavefunction(dev["out"], dev["out"].download())
    assert _relerr(got, ref) < RTOL


def test_heff_plan_outlives_block_hamiltonians(ctx):
    # the plan keeps its own copies of H_L and H_R: freeing the caller's
    # tensors (and reusing their memory) must not change later applies
    from su2dmrg_amd import lib as L
    _say("heff plan after hl/hr are destroyed")
    work = synthetic.small(seed=5, twoj=(1, 3, 5), dims=(60, 90, 20), nterm=2)
    dev = L.upload_workload(ctx, work)
    dev["hl"].close()
    dev["hr"].close()
    junk = [L.scalar_operator(ctx, work["twoj"], work["dims_a"],
                              {j: np.full((d, d), 7.0) for j, d in zip(work["twoj"],
                                                                      work["dims_a"])})
            for _ in range(2)]
    dev["plan"].apply(dev["psi"], dev["out"])
    got = L.unpack_wavefunction(dev["out"], dev["out"].download())
    ref = oheff.heff_apply(work["psi"], work["hl"], work["hr"], work["terms"])
    assert _relerr(got, ref) < RTOL
    for j in junk:
        j.close()


# FP32 MFMA matvec (FP32 accumulation over K up to ~25000 terms): a
# relative 2-norm error of a few 1e-6 is expected (rounding of the operands
# to FP32, 2.4e-8 relative rms, plus accumulation error growing with
# sqrt(K / 4)); the bound is 5e-5.  The lower bound shows the FP32 path ran.
RTOL_FP32 = 5e-5


@pytest.mark.parametrize("seed,twoj,dims,nterm", [
    (0, (0, 1, 2, 3, 4), (3, 4, 5, 2, 1), 3),
    (2, (0, 2, 4, 6), (130, 257, 129, 7), 4),
])
def test_heff_fp32_small_matches_oracle(ctx, seed, twoj, dims, nterm):
    from su2dmrg_amd import lib as L
    _say(f"heff fp32 seed={seed} twoj={twoj} dims={dims}")
    work = synthetic.small(seed=seed, twoj=twoj, dims=dims, nterm=nterm)
    dev = L.upload_workload(ctx, work)
    dev["plan"].set_precision(32)
    dev["plan"].apply(dev["psi"], dev["out"])
    got = L.unpack_wavefunction(dev["out"], dev["out"].download())
    ref = oheff.heff_apply(work["psi"], work["hl"], work["hr"], work["terms"])
    err = _relerr(got, ref)
    _say(f"  fp32 rel err {err:.3e}")
    assert 1e-10 < err < RTOL_FP32
    # back to FP64 on the same plan
    dev["plan"].set_precision(64)
    dev["plan"].apply(dev["psi"], dev["out"])
    got = L.unpack_wavefunction(dev["out"], dev["out"].download())
    assert _relerr(got, ref) < RTOL


def test_heff_fp32_cyl16x8_d2000_every_variant(ctx, headline):
    from su2dmrg_amd import lib as L
    _say("heff fp32 cyl16x8_d2000 seed=0, every tile variant")
    work, ref = headline
    dev = L.upload_workload(ctx, work)
    plan = dev["plan"]
    plan.set_precision(32)
    chosen, t1, t2 = plan.autotune(reps=1)
    _say(f"  fp32 autotune {chosen} stage1 {['%.2f' % t for t in t1]} "
         f"stage2 {['%.2f' % t for t in t2]}")
    nv = len(t1)
    for v in range(nv):
        plan.set_variants(v, v)
        plan.apply(dev["psi"], dev["out"])
        got = L.unpack_wavefunction(dev["out"], dev["out"].download())
        err = _relerr(got, ref)
        _say(f"  variant {v}: fp32 rel err {err:.3e}")
        assert 1e-10 < err < RTOL_FP32
