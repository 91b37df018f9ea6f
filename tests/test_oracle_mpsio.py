"""Checkpoint files in the reference's format, oracle side (CPU only)."""
import os

import numpy as np
import pytest

from oracle import dmrg as od
from oracle import mpsio
from oracle.su2 import block_index, cgc, valid_blocks


@pytest.fixture(scope="module")
def saved(tmp_path_factory):
    # 4 (open) x 2 (open) J1 = J2 = 1 at D = 100: one sweep, then a checkpoint
    lat = od.Lattice(4, 2, jcoup=1.0, bdry=0)
    dm = od.DMRG(lat, totspin=0, max_dcut=100)
    dm.run(nsweep=1, stop=None)
    d = str(tmp_path_factory.mktemp("ckpt"))
    mpsio.save(dm, d)
    return lat, dm, d


def test_file_names_and_header(saved):
    lat, dm, d = saved
    names = sorted(os.listdir(d))
    assert "enr-8-100-0-0-0.dat" in names and "ww-8-100-0-0-0-4.bin" in names
    assert all(f"uu-8-100-0-0-0-{i}.{e}" in names for i in range(8) for e in ("dat", "bin"))
    # site 3 (left of the centre): [lb[2], site, lb[3]], directions 1, 1, -1
    lines = open(os.path.join(d, "uu-8-100-0-0-0-3.dat")).read().split("\n")
    assert lines[0] == "3\t0"
    assert lines[1] == f"1\t{dm.lb[2].nsec}"
    assert lines[2] == "".join(f"{j}\t" for j in dm.lb[2].twoj)
    assert lines[3] == "".join(f"{x}\t" for x in dm.lb[2].dims)
    assert lines[4:7] == ["1\t1", "1\t", "1\t"]
    assert lines[7] == f"-1\t{dm.lb[3].nsec}"
    # the energy line is the Ritz value minus 1000 (save_enr)
    e = float(open(os.path.join(d, "enr-8-100-0-0-0.dat")).read().split()[-1])
    assert abs(e + 1000 - dm.energies[-1]) < 1e-12


def test_binary_body_layout(saved):
    lat, dm, d = saved
    legs, blocks = mpsio.read_tensor3(os.path.join(d, "uu-8-100-0-0-0-3"))
    raw = open(os.path.join(d, "uu-8-100-0-0-0-3.bin"), "rb").read()
    nten = legs[0].nsec * legs[1].nsec * legs[2].nsec
    flags = np.frombuffer(raw[:4 * nten], dtype=np.int32)
    # exactly the allowed blocks whose outgoing sector is in the bond
    want = [labs for labs in valid_blocks(legs, 0) if labs[2] in dm.lb[3].twoj]
    assert sorted(block_index(legs, labs) for labs in want) == list(np.flatnonzero(flags))
    body = np.frombuffer(raw[4 * nten:], dtype=np.float64)
    nred = sum(blocks[labs].size for labs in want)
    cg = np.concatenate([np.asarray(cgc(*labs)).ravel(order="F") for labs in want])
    np.testing.assert_array_equal(body[nred:], cg)
    # reduced block (ja, 1, jb) = projector rows of the part ja in sector jb
    E = od.Enlarged(dm.lb[2], left=True)
    for (ja, s, jb) in want:
        o = E.off[(ja, jb)]
        d_a = dm.lb[2].dims[dm.lb[2].index(ja)]
        np.testing.assert_array_equal(blocks[(ja, s, jb)][:, 0, :], dm.U[3][jb][o:o + d_a])


def test_round_trip_continues_the_same_trajectory(saved, tmp_path):
    lat, dm, d = saved
    ref = od.DMRG(lat, totspin=0, max_dcut=100)
    ref.run(nsweep=2, stop=None)
    re = od.DMRG(lat, totspin=0, max_dcut=100)
    e0 = mpsio.load(re, d)
    assert abs(e0 - dm.energies[-1]) < 1e-12
    re.sweep()
    n = len(re.energies)
    np.testing.assert_array_equal(re.energies, ref.energies[-n:])


def test_scaled_cg_is_folded_and_a_foreign_cg_is_rejected(saved, tmp_path):
    lat, dm, d = saved
    legs, blocks = mpsio.read_tensor3(os.path.join(d, "uu-8-100-0-0-0-3"))
    path = str(tmp_path / "t")
    mpsio.write_tensor3(path, legs, blocks)
    raw = bytearray(open(path + ".bin", "rb").read())
    nten = legs[0].nsec * legs[1].nsec * legs[2].nsec
    nred = sum(b.size for b in blocks.values())
    body = np.frombuffer(bytes(raw[4 * nten:]), dtype=np.float64).copy()
    first = sorted(blocks, key=lambda labs: block_index(legs, labs))[0]
    ncg0 = np.asarray(cgc(*first)).size
    # a CG stored at twice make_cgc: the reader folds the 2 into the block
    b2 = body.copy()
    b2[nred:nred + ncg0] *= 2.0
    open(path + ".bin", "wb").write(bytes(raw[:4 * nten]) + b2.tobytes())
    _, blk2 = mpsio.read_tensor3(path)
    np.testing.assert_allclose(blk2[first], 2.0 * blocks[first], rtol=1e-14)
    # a CG that is no multiple of make_cgc is an error
    b3 = body.copy()
    b3[nred] += 0.5
    open(path + ".bin", "wb").write(bytes(raw[:4 * nten]) + b3.tobytes())
    with pytest.raises(ValueError):
        mpsio.read_tensor3(path)


def test_wout_lists_kept_then_cut_values():
    rng = np.random.default_rng(3)
    psi = {1: rng.standard_normal((6, 5)), 3: rng.standard_normal((4, 4))}
    info = {}
    U, V, bond, kept = od.truncated_svd(psi, 1, 0, 6, info)
    w = info["wout"]
    s1 = np.sort(np.linalg.svd(psi[1], compute_uv=False))
    s3 = np.sort(np.linalg.svd(psi[3], compute_uv=False))
    k1, k3 = bond.dims[bond.index(1)] if 1 in bond.twoj else 0, \
        bond.dims[bond.index(3)] if 3 in bond.twoj else 0
    want = np.concatenate([s1[len(s1) - k1:], s3[len(s3) - k3:], s1[:len(s1) - k1],
                           s3[:len(s3) - k3]])
    np.testing.assert_allclose(w, want, rtol=1e-10)
    assert k1 + k3 == 6
