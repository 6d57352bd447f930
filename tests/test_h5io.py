"""HDF5 readers and writers of the reference's file formats (adga_amd/h5io.py over libadga_h5.so).

Pins, none of which restates the library's own code:
* the reference's templates (documentation/hdf5_templates, copied to tests/golden/): the 1P reader takes the
  real layout (dmft-last/ineq-001/siw|giw/value: 3 bands x 2 spins x 100 frequencies, dc, mu, the beta
  attribute) and the V(q) reader the real V(q) layout (.axes/Q-points, one dataset per band pattern);
* files written with the generic dataset writer (no knowledge of the layouts) whose entries encode their own
  coordinates, so that a transposed or shifted index shows as a wrong decoded value, checked against the
  reference's statements of each layout (read_siw/giw spin average and atom placement, read_vq's compound
  index, read_vertex's compound index);
* closed forms of read_vertex's disconnected terms (vertex_type 0, 1, 2) with one and two atoms;
* h5dump (the HDF5 distribution's own tool) for the datasets the writers produce: dims, filters and values.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT

h5io = pytest.importorskip("adga_amd.h5io", reason="libadga_h5.so not built (needs the HDF5 C library)")
GOLD = os.path.join(ROOT, "tests", "golden")
H5DUMP = shutil.which("h5dump") or "/opt/conda/bin/h5dump"


def _dump(path, dset, extra=()):
    if not os.path.exists(H5DUMP):
        pytest.skip("h5dump not available")
    return subprocess.run([H5DUMP, *extra, "-d", dset, path], capture_output=True, text=True, check=True).stdout


def _dump_values(path, dset):
    """Values of a compound (r, i) dataset as h5dump prints them, in file order."""
    txt = _dump(path, dset)
    body = txt[txt.index("DATA {"):]
    nums = []
    for line in body.splitlines():
        line = line.strip().rstrip(",")
        if not line or line.startswith("(") or line in ("{", "}", "DATA {"):
            continue
        try:
            nums.append(float(line))
        except ValueError:
            pass
    v = np.array(nums)
    return v[0::2] + 1j * v[1::2]


# ---------------------------------------------------------------------- 1P file
def test_1p_template_layout():
    with h5io.H5File(os.path.join(GOLD, "DMFT_1PG_template.hdf5")) as f:
        assert f.dims("dmft-last/ineq-001/siw/value") == (3, 2, 100)
        siw = f.read_siw("dmft-last", 3, [[3, 0]], 50)
        giw = f.read_giw("dmft-last", 3, [[3, 0]], 50)
        dc = f.read_dc("dmft-last", 3, [[3, 0]])
        assert siw.shape == (100, 3) and giw.shape == (100, 3) and dc.shape == (2, 3)
        assert f.read_beta() == 10.0
        assert f.read_mu("dmft-last") == 0.0
        with pytest.raises(h5io.H5Error):  # the file holds 2*50 frequencies, not 2*40
            f.read_siw("dmft-last", 3, [[3, 0]], 40)


def _encoded_1p(path, atoms, iwmax, nk=(2, 2, 1)):
    """A 1P file with entries that encode their coordinates: value(b, s, w) = 1000 a + 100 b + 10 s + i w
    on atom a."""
    n = sum(c + u for c, u in atoms)
    with h5io.H5File(path, "w") as f:
        for a, (c, _) in enumerate(atoms, start=1):
            b, s, w = np.meshgrid(np.arange(c), np.arange(2), np.arange(2 * iwmax), indexing="ij")
            for name, extra in (("siw", 0.0), ("giw", 0.5)):
                f.write_complex(f"dmft-last/ineq-{a:03d}/{name}/value", 1000 * a + 100 * b + 10 * s + extra + 1j * w)
            f.write_real(f"dmft-last/ineq-{a:03d}/dc/value", 10.0 * a + np.arange(c)[:, None] + 0.1 * np.arange(2)[None, :])
        f.write_real("dmft-last/mu/value", np.array(1.25))
        f.write_attr(".config", "general.beta", 7.5)
        nkp = int(np.prod(nk))
        hk = (np.arange(n)[:, None, None] + 10 * np.arange(n)[None, :, None] + 100 * np.arange(nkp)[None, None, :]
              + 1j * np.arange(nkp)[None, None, :])
        f.write_complex("start/hk/value", np.ascontiguousarray(hk.transpose(2, 1, 0)))  # Fortran hk(n, n, nkp)
    return n, hk


@pytest.mark.parametrize("orb_sym", [False, True])
def test_1p_reader_decodes_orientation_and_atoms(tmp_path, orb_sym):
    atoms = [(2, 1), (1, 0)]  # atom 1: two correlated + one uncorrelated orbital, atom 2: one correlated
    iwmax = 4
    path = str(tmp_path / "p1.hdf5")
    n, hk = _encoded_1p(path, atoms, iwmax)
    with h5io.H5File(path) as f:
        siw = f.read_siw("dmft-last", n, atoms, iwmax, orb_sym)
        giw = f.read_giw("dmft-last", n, atoms, iwmax, orb_sym)
        dc = f.read_dc("dmft-last", n, atoms)
        assert f.read_mu("dmft-last") == 1.25 and f.read_beta() == 7.5
        assert np.array_equal(f.read_hk(), hk)
    w = np.arange(2 * iwmax)
    # read_siw: spin average on the correlated orbitals, frequency row iw + iwmax, zero elsewhere
    expect = {0: 1000 + 5 + 1j * w, 1: 1100 + 5 + 1j * w, 3: 2000 + 5 + 1j * w}
    if orb_sym:
        expect[0] = expect[1] = 1050 + 5 + 1j * w
    for orb in range(n):
        e = expect.get(orb, 0.0 * w)
        assert np.array_equal(siw[:, orb], e), orb
        assert np.array_equal(giw[:, orb], e + (0.5 if orb in expect else 0.0)), orb
    assert np.array_equal(dc.real, np.array([[10, 11, 0, 20], [10.1, 11.1, 0, 20.1]]))


# ---------------------------------------------------------------------- V(q) file
def test_vq_template_layout():
    with h5io.H5File(os.path.join(GOLD, "VQ_template.hdf5")) as f:
        v = f.read_vq(1, 8000)
        assert v.shape == (1, 1, 8000)
        with pytest.raises(h5io.H5Error):  # the q volume of the file has 8000 points
            f.read_vq(1, 64)


def test_vq_reader_compound_index(tmp_path):
    n, nqp = 2, 6
    path = str(tmp_path / "vq.hdf5")
    with h5io.H5File(path, "w") as f:
        f.write_real(".axes/Q-points", np.zeros((nqp, 3)))
        for ind in range(1, n ** 4 + 1):
            f.write_complex(f"{ind:05d}", ind + 1j * np.arange(nqp))
    with h5io.H5File(path) as f:
        v = f.read_vq(n, nqp)
        vp = f.read_vq(n, 2, q_vol=False, q_data=[5, 2])
    for i in range(n):
        for j in range(n):
            for k in range(n):
                for l in range(n):
                    ind = n ** 3 * i + n ** 2 * j + n * k + l + 1     # component2index_band, 0-based bands
                    i1, i2 = i * n + k, l * n + j                     # read_vq: v({k i}, {j l}) = vq(i, j, k, l)
                    assert np.array_equal(v[i1, i2, :], ind + 1j * np.arange(nqp))
                    assert np.array_equal(vp[i1, i2, :], ind + 1j * np.array([4, 1]))


# ---------------------------------------------------------------------- 2P (vertex) file
def _g4_encoded(nb, iwfmax, channel):
    """g4(b1, b2, nu1, b3, b4, nu2) whose value encodes its coordinates."""
    b1, b2, u1, b3, b4, u2 = np.meshgrid(*[np.arange(x) for x in (nb, nb, 2 * iwfmax, nb, nb, 2 * iwfmax)],
                                         indexing="ij")
    return np.asfortranarray(b1 + 10 * b2 + 100 * b3 + 1000 * b4 + 1j * (u1 + 100 * u2 + 10000 * channel))


def _giw(n, iwmax, seed=3):
    rng = np.random.default_rng(seed)
    return np.asfortranarray(rng.normal(size=(2 * iwmax, n)) + 1j * rng.normal(size=(2 * iwmax, n)))


def test_vertex_reader_compound_index(tmp_path):
    """vertex_type 1 with g4 encoding its coordinates: chi(i1 = {nu1, i, j}, i2 = {nu2, l, k}) =
    beta g4(i, j, nu1, k, l, nu2) minus beta g_i(nu1) g_j(nu1 - omega) where nu2 = nu1, l = i, k = j."""
    n, iwfmax, iwf, iwbmax, iwmax, beta = 2, 3, 2, 1, 6, 4.0
    path = str(tmp_path / "v.hdf5")
    with h5io.H5File(path, "w") as f:
        for iwb in range(-iwbmax, iwbmax + 1):
            f.write_vertex(1, iwb, iwbmax, _g4_encoded(n, iwfmax, 1), _g4_encoded(n, iwfmax, 0))
    g = _giw(n, iwmax)
    with h5io.H5File(path) as f:
        for iwb in (-1, 0, 1):
            d, m = f.read_vertex(iwb, n, [[n, 0]], iwfmax, iwbmax, iwf, beta, g, vertex_type=1)
            for ch, x in ((1, d), (0, m)):
                for w1 in range(-iwf, iwf):
                    for i in range(n):
                        for j in range(n):
                            for w2 in range(-iwf, iwf):
                                for l in range(n):
                                    for k in range(n):
                                        i1 = (w1 + iwf) * n * n + i * n + j
                                        i2 = (w2 + iwf) * n * n + l * n + k
                                        u1, u2 = w1 + iwfmax, w2 + iwfmax
                                        e = beta * (i + 10 * j + 100 * k + 1000 * l + 1j * (u1 + 100 * u2 + 10000 * ch))
                                        if w1 == w2 and i == l and j == k:
                                            e -= beta * g[w1 + iwmax, i] * g[w2 - iwb + iwmax, j]
                                        assert x[i1, i2] == pytest.approx(e, rel=1e-15, abs=1e-12)


@pytest.mark.parametrize("vertex_type", [0, 1, 2])
def test_vertex_disconnected_terms(tmp_path, vertex_type):
    """g4 = 0 on two atoms (orbitals 0-1 on atom 1, orbital 2 on atom 2): what remains are the disconnected
    terms of read_vertex alone (hdf5_module.f90:729-763): type 0 subtracts 2 beta g_i g_l in the density
    channel at omega = 0 for i = j, k = l on one atom, and the bubble where the legs are not all on one
    atom; type 1 the bubble everywhere; type 2 the bubble where the legs are not all on one atom."""
    n, iwfmax, iwf, iwbmax, iwmax, beta = 3, 2, 2, 1, 6, 5.0
    atoms = [[2, 0], [1, 0]]
    path = str(tmp_path / "v2.hdf5")
    with h5io.H5File(path, "w") as f:
        for iwb in range(-iwbmax, iwbmax + 1):
            z2 = np.zeros((2, 2, 2 * iwfmax, 2, 2, 2 * iwfmax), order="F")
            z1 = np.zeros((1, 1, 2 * iwfmax, 1, 1, 2 * iwfmax), order="F")
            f.write_vertex(1, iwb, iwbmax, z2, z2)
            f.write_vertex(2, iwb, iwbmax, z1, z1)
    atom = [1, 1, 2]
    g = _giw(n, iwmax, seed=5)
    with h5io.H5File(path) as f:
        for iwb in (0, 1):
            d, m = f.read_vertex(iwb, n, atoms, iwfmax, iwbmax, iwf, beta, g, vertex_type)
            N = n * n * 2 * iwf
            ed = np.zeros((N, N), dtype=complex)
            em = np.zeros((N, N), dtype=complex)
            for w1 in range(-iwf, iwf):
                for w2 in range(-iwf, iwf):
                    for i in range(n):
                        for j in range(n):
                            for l in range(n):
                                for k in range(n):
                                    i1 = (w1 + iwf) * n * n + i * n + j
                                    i2 = (w2 + iwf) * n * n + l * n + k
                                    same = atom[i] == atom[j] == atom[k] == atom[l]
                                    bub = beta * g[w1 + iwmax, i] * g[w2 - iwb + iwmax, j]
                                    is_bub = w1 == w2 and i == l and j == k
                                    if vertex_type == 0:
                                        if iwb == 0 and i == j and k == l and same:
                                            ed[i1, i2] -= 2 * beta * g[w1 + iwmax, i] * g[w2 + iwmax, l]
                                        if is_bub and not same:
                                            ed[i1, i2] -= bub
                                            em[i1, i2] -= bub
                                    elif vertex_type == 1 or not same:
                                        if is_bub:
                                            ed[i1, i2] -= bub
                                            em[i1, i2] -= bub
            assert np.allclose(d, ed, rtol=1e-15, atol=1e-13) and np.allclose(m, em, rtol=1e-15, atol=1e-13)


# ---------------------------------------------------------------------- output file
def _chi_random(n, npts, seed=1):
    rng = np.random.default_rng(seed)
    return np.asfortranarray(rng.normal(size=(n * n, n * n, npts)) + 1j * rng.normal(size=(n * n, n * n, npts)))


def test_output_file_layouts(tmp_path):
    n, iwbm, nq, iwf = 2, 1, (2, 1, 3), 3
    nw, nqp = 2 * iwbm + 1, int(np.prod(nq))
    path = str(tmp_path / "out.hdf5")
    chi = _chi_random(n, nqp * nw)
    chil = _chi_random(n, nw, seed=2)
    rng = np.random.default_rng(4)
    nkp = nqp
    ss = np.asfortranarray(rng.normal(size=(n, n, 2 * iwf, nkp)) + 1j * rng.normal(size=(n, n, 2 * iwf, nkp)))
    shf = np.asfortranarray(rng.normal(size=(n, n, nkp)) + 0j)
    sloc = np.asfortranarray(rng.normal(size=(n, n, 2 * iwf)) + 1j)
    sd = np.asfortranarray(rng.normal(size=(n, n, 2 * iwf)) - 1j)
    ndk = np.asfortranarray(rng.normal(size=(nkp, n, n)) + 0j)
    hk = np.asfortranarray(rng.normal(size=(n, n, nkp)) + 1j * rng.normal(size=(n, n, nkp)))
    with h5io.H5File(path, "w") as f:
        f.init_output(ndim=n, iwmax=8, iwfmax=4, iwbmax=2, iwfmax_small=iwf, iwbmax_small=iwbm, nkp=nkp,
                      nkpx=nq[0], nkpy=nq[1], nkpz=nq[2], nqp=nqp, nqpx=nq[0], nqpy=nq[1], nqpz=nq[2], nkp_eom=nkp,
                      do_chi=1, do_eom=1, mu=0.5, beta=9.0, hk=hk, n_dmft=np.array([0.5, 0.25]), **{"nonlocal": 1})
        f.write_chi_qw("magn", chi, n, iwbm, nq, full=True)
        f.write_chi_qw("dens", chi, n, iwbm, nq, full=False)
        f.write_chi_loc("magn", chil, n, iwbm, full=True)
        f.write_chi_loc("dens", chil, n, iwbm, full=False)
        f.write_eom(sd, n, iwf, nq, sigma_sum=ss, sigma_sum_hf=shf, sigma_loc=sloc)
        f.write_occ(np.array([0.4, 0.6]), ndk, n, nq)
    with h5io.H5File(path) as f:
        assert f.read_real("input/mu") == 0.5 and f.read_real("input/beta") == 9.0
        h = f.read_complex("input/hk")  # hk_arr(kz, ky, kx, i2, i1) = hk(i1, i2, k)
        for i1 in range(n):
            for i2 in range(n):
                for kx in range(nq[0]):
                    for ky in range(nq[1]):
                        for kz in range(nq[2]):
                            assert h[i1, i2, kx, ky, kz] == hk[i1, i2, kz + ky * nq[2] + kx * nq[1] * nq[2]]
        full = f.read_complex("susceptibility/nonloc/magn")
        red = f.read_complex("susceptibility/nonloc/dens")
        assert full.shape == (n, n, n, n, nq[0], nq[1], nq[2], nw) and red.shape == (n, n, nq[0], nq[1], nq[2], nw)
        # output_chi_qw_full_h5: chi_slice(qz, qy, qx, i4, i3, i2, i1) = chi({i1 i2}, {i4 i3}, point)
        for (i1, i2, i3, i4, qx, qy, qz, w) in np.ndindex(full.shape):
            p = w * nqp + qz + qy * nq[2] + qx * nq[1] * nq[2]
            assert full[i1, i2, i3, i4, qx, qy, qz, w] == chi[i1 * n + i2, i4 * n + i3, p]
            if i2 == i1 and i4 == i3:
                assert red[i1, i3, qx, qy, qz, w] == chi[i1 * n + i1, i3 * n + i3, p]
        cl = f.read_complex("susceptibility/loc/magn")
        for (i1, i2, i3, i4, w) in np.ndindex(cl.shape):
            assert cl[i1, i2, i3, i4, w] == chil[i1 * n + i2, i4 * n + i3, w]
        sk = f.read_complex("selfenergy/nonloc/dga")
        hf = f.read_complex("selfenergy/nonloc/hartree_fock")
        for (i1, i2, kx, ky, kz, w) in np.ndindex(sk.shape):
            k = kz + ky * nq[2] + kx * nq[1] * nq[2]
            assert sk[i1, i2, kx, ky, kz, w] == ss[i1, i2, w, k] and hf[i1, i2, kx, ky, kz, w] == shf[i1, i2, k]
        assert np.array_equal(f.read_complex("selfenergy/loc/dga_ksum"), sloc.transpose(0, 1, 2))
        assert np.array_equal(f.read_complex("selfenergy/loc/dmft"), sd)
        assert np.array_equal(f.read_real("occupation/n_dga"), [0.4, 0.6])
        nk5 = f.read_real("occupation/n_dga_k")
        for (i1, i2, kx, ky, kz) in np.ndindex(nk5.shape):
            assert nk5[i1, i2, kx, ky, kz] == ndk[kz + ky * nq[2] + kx * nq[1] * nq[2], i1, i2].real
    # HDF5's own tool sees the reference's filters and chunks, and the same values
    hdr = _dump(path, "/susceptibility/nonloc/magn", ("-p", "-H"))
    for word in ("DEFLATE", "SHUFFLE", "FLETCHER32", f"CHUNKED ( {n}, {n}, {n}, {n}, {nq[0]}, {nq[1]}, {nq[2]}, 1 )"):
        assert word in hdr, word
    vals = _dump_values(path, "/susceptibility/loc/dens")
    expect = np.array([chil[i1 * n + i1, i3 * n + i3, w] for i1 in range(n) for i3 in range(n) for w in range(nw)])
    assert np.allclose(vals, expect, rtol=1e-5, atol=1e-6)  # h5dump prints 6 significant digits


def test_chi_qpath_layouts(tmp_path):
    n, iwbm, nqp = 2, 1, 5
    nw = 2 * iwbm + 1
    chi = _chi_random(n, nqp * nw, seed=7)
    path = str(tmp_path / "qp.hdf5")
    with h5io.H5File(path, "w") as f:
        f.init_output(ndim=n, iwmax=4, do_chi=1, nqp=nqp, mu=0.0, beta=1.0, **{"nonlocal": 1})
        f.write_chi_qpath("magn", chi, n, iwbm, nqp, full=True)
        f.write_chi_qpath("dens", chi, n, iwbm, nqp, full=False)
    with h5io.H5File(path) as f:
        full = f.read_complex("susceptibility/nonloc/magn")
        red = f.read_complex("susceptibility/nonloc/dens")
    # output_chi_qpath_full_h5: (w, q, i4, i3, i2, i1); reduced: (w, q, i1, i3)
    for (i1, i2, i3, i4, q, w) in np.ndindex(full.shape):
        assert full[i1, i2, i3, i4, q, w] == chi[i1 * n + i2, i4 * n + i3, w * nqp + q]
    for (i3, i1, q, w) in np.ndindex(red.shape):
        assert red[i3, i1, q, w] == chi[i1 * n + i1, i3 * n + i3, w * nqp + q]
