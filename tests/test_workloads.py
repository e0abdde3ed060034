"""The synthetic states the benchmark builds (CPU only).

The 512x256x256 state is never built whole: its pieces must add up to the
record a one-shot build would give.  Checked on the first pieces against
the quiet-start placement of _quiet_cells, the density profile and ids."""
import numpy as np

from mandor3_amd import workloads


def test_large_pieces_follow_the_quiet_start_and_profile():
    wl = workloads.large_foil(planes_per_chunk=1)
    n, h = wl["n"], wl["h"]
    assert wl["n_markers"] == 2 * 16 * (n[0] - 2) * (n[1] - 2) * (n[2] - 2)
    assert wl["n_markers"] < 2 ** 31
    gen = wl["chunks"]()
    nid = 1
    for i in (1, 2):                                   # planes 1 and 2, one piece each
        M = next(gen).copy()
        cells = np.stack(np.meshgrid([i], np.arange(1, n[1] - 1), np.arange(1, n[2] - 1),
                                     indexing="ij"), -1).reshape(-1, 3).astype(np.float64)
        pos, ppc = workloads._quiet_cells(cells, (4, 2, 2), h, None)
        N = pos.shape[0]
        assert ppc == 16 and M.size == 2 * N
        for s in range(2):
            sl = slice(s * N, (s + 1) * N)
            assert np.array_equal(M["x"][sl], pos[:, 0])
            assert np.array_equal(M["y"][sl], pos[:, 1])
            assert np.array_equal(M["z"][sl], pos[:, 2])
        # 0.1 n_c floor far from the foil, electrons negative, cold protons
        w = 0.1 * workloads.RHO_CRIT / 16
        assert np.allclose(M["rho"][:N], -w) and np.allclose(M["rho"][N:], w)
        assert np.all(M["qDivM"][:N] == -1.0) and np.all(M["vx"][N:] == 0.0)
        assert 0.009 < M["vx"][:N].std() < 0.011
        assert np.array_equal(M["id"], np.arange(nid, nid + 2 * N))
        nid += 2 * N
    assert wl["plasma_cells"] == (n[0] - 2) * (n[1] - 2) * (n[2] - 2)
    # profile along x: 40 n_c foil on planes [256, 280), exponential
    # preplasma (scale 0.5 lambda) on the 48 planes before it, 0.1 n_c floor
    d = wl["density"]
    assert d(256) == 40.0 and d(279) == 40.0 and d(280) == 0.1 and d(207) == 0.1
    assert np.isclose(d(255), 40.0 * np.exp(-0.5 * h[0] / 0.5))
    assert np.isclose(d(208), 40.0 * np.exp(-47.5 * h[0] / 0.5))
    assert d(210) > d(208) > 0.1
