"""The oracle's restatement of the reference CPU path (compat semantics) against
fixtures written by the compiled reference (tests/golden/make_golden.py).
Every comparison is bit-exact: same inputs, same float64/float32 results."""
import numpy as np


def bits_equal(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return a.shape == b.shape and a.tobytes() == b.tobytes()


def test_gauss_constants(orc, compat_cases):
    d1, d2 = orc.gauss_consts(1.0, 0.55)
    assert bits_equal(np.array([d1, d2]), compat_cases[0]["gauss"])
    # closed form of eq. 6.8 at resolution 1, outlier 0.55
    c1, c2 = 10 * (1 - 0.55), 0.55
    d3 = -np.log(c2)
    assert np.isclose(d1, -np.log(c1 + c2) - d3, rtol=0, atol=1e-15)
    assert np.isclose(d2, -2 * np.log((-np.log(c1 * np.exp(-0.5) + c2) - d3) / d1), rtol=0, atol=1e-15)


def test_grid_build(orc, compat_cases):
    for c in compat_cases[::3]:
        g = orc.CompatGrid(c["map"])
        assert bits_equal(np.array(g.g.minv), c["min"]) and bits_equal(np.array(g.g.maxv), c["max"])
        assert tuple(g.dims) == tuple(c["dims"]) and g.g.ncells == int(c["ncells"][0])
        idx, n, mean, icov = g.occupied()
        assert bits_equal(idx, c["cell_idx"]) and bits_equal(n.astype(np.int32), c["cell_n"])
        assert bits_equal(mean, c["cell_mean"]) and bits_equal(icov, c["cell_icov"])


def test_radius_search(orc, compat_cases):
    for c in compat_cases:
        g = orc.CompatGrid(c["map"])
        scan = np.concatenate([c["scan"][:, :3], np.ones((c["scan"].shape[0], 1), np.float32)], 1)
        trans = orc.compat_transform(scan, c["guess"])
        _, _, mean, _ = g.occupied()
        idx_all, _, _, _ = g.occupied()
        pos = {int(i): k for k, i in enumerate(idx_all)}
        off = 0
        for i, cnt in enumerate(c["radius_count"]):
            cells, _ = g.radius_search(trans[i])
            assert len(cells) == cnt
            got = np.array([mean[pos[int(ci)]] for ci in cells]).reshape(-1, 3)
            assert bits_equal(got, c["radius_mean"][off:off + cnt])
            off += cnt


def test_first_derivative_pass(orc, compat_cases):
    d1, d2 = orc.gauss_consts()
    for c in compat_cases:
        g = orc.CompatGrid(c["map"])
        trans = orc.compat_transform(c["scan"], c["guess"])
        p0 = np.r_[c["guess"][:3, 3].astype(np.float64), orc.compat_euler(c["guess"])]
        assert bits_equal(p0, c["p0"])
        s, gr, H, _ = g.derivatives(c["scan"], trans, p0, d1, d2)
        assert bits_equal(np.array([s]), c["score0"])
        assert bits_equal(gr, c["grad0"]) and bits_equal(H, c["hess0"])


def test_end_to_end(orc, compat_cases):
    for c in compat_cases:
        for omp in (False, True):
            r = orc.compat_align(c["map"], c["scan"], c["guess"], omp=omp)
            assert bits_equal(r["final_f32"], c["final"])
            assert int(r["converged"]) == int(c["converged"][0]) and r["iterations"] == int(c["iterations"][0])
            # the reference's invariants: 2 iterations, converged, translation = guess
            assert r["iterations"] == 2 and r["converged"]
            assert bits_equal(r["final_f32"][:3, 3], c["guess"][:3, 3])


def test_math_functions(orc, compat_math):
    m = compat_math
    d1, d2 = -2.2172252440428890, 0.43312300470355464
    for i in range(m["p"].shape[0]):
        j, h = orc.angle_tables(m["p"][i])
        assert bits_equal(j.ravel(), m["out_jang"][i]) and bits_equal(h.ravel(), m["out_hang"][i])
        pg, ph = orc.point_derivs(m["x"][i], m["p"][i])
        assert bits_equal(pg.ravel(), m["out_pg"][i]) and bits_equal(ph.ravel(), m["out_ph"][i])
        assert bits_equal(orc.solve6(m["A"][i], m["b"][i]), m["out_solve"][i])
        mt = m["mt"][i]
        assert bits_equal(np.array([orc.mt_trial(*mt)]), np.array([m["out_trial"][i]]))
        ret, vals = orc.mt_update(*mt)
        assert ret == int(m["out_upd_ret"][i]) and bits_equal(np.array(vals), m["out_upd"][i])
        inc, g6, h6 = orc.compat_update(m["xt"][i], m["cinv"][i], m["x"][i], m["p"][i], d1, d2)
        assert bits_equal(np.array([inc]), np.array([m["out_inc"][i]]))
        assert bits_equal(g6, m["out_grad"][i]) and bits_equal(h6.ravel(), m["out_hess"][i])
        B = orc.compat_build_matrix(m["M"][i], m["p"][i])
        assert bits_equal(B.ravel(), m["out_build"][i])
        assert bits_equal(orc.compat_euler(m["M"][i]), m["out_euler"][i])


def test_oracle_pins_catch_planted_mistakes(orc, compat_math):
    """A plausible slip in the oracle (dropped term, swapped index, sign) must fail
    the fixtures above: perturb the oracle's outputs the way such slips would and
    check that the fixtures tell them apart."""
    m = compat_math
    i = 7
    pg, ph = orc.point_derivs(m["x"][i], m["p"][i])
    assert not bits_equal(pg.T.ravel(), m["out_pg"][i])           # transposed operand
    sol = orc.solve6(m["A"][i], m["b"][i])
    assert not bits_equal(-sol, m["out_solve"][i])                 # sign
    inc, g6, h6 = orc.compat_update(m["xt"][i], m["cinv"][i], m["x"][i], m["p"][i], -2.2172252440428890,
                                    0.43312300470355464)
    assert not bits_equal(h6.T.ravel(), m["out_hess"][i])          # the quirky Hessian is not symmetric
