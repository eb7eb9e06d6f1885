"""The oracle's More-Thuente line search (computeStepLengthMT, reference
src/CPU/ndt_mapping/kernel.cpp:1015-1144, with the loop live) on scalar functions
whose steps follow from closed forms.  The interpolation helpers it calls are pinned
bit for bit against the compiled reference (tests/test_oracle_compat.py); these
tests pin the loop around them: which trial it takes, when it stops, the direction
reversal, the clamps and the evaluation sequence the NDT driver relies on (the last
evaluation is at the returned step).  The reference's own start
(interval_converged = (step_max - step_min) > 0, :1053, a loop that never runs) is
planted through a flag and must fail."""
import math

import numpy as np
import pytest

MU, NU = 1e-4, 0.9


def cubic(t):
    """phi(t) = t^3 / 3 - t: phi'(0) = -1, minimum at t = 1."""
    return t ** 3 / 3 - t, t * t - 1


def quadratic(c, t_star):
    return lambda t: (0.5 * c * (t - t_star) ** 2, c * (t - t_star))


def wolfe(phi, a, sign, phi_0, d_phi_0):
    """The loop's acceptance test, evaluated independently: sufficient decrease
    psi(a) <= 0 and phi'(a) <= -nu phi'(0) along the descent direction."""
    v, d = phi(sign * a)
    d_s, d0 = sign * d, -abs(d_phi_0)
    return v - phi_0 - MU * d0 * a <= 0 and d_s <= -NU * d0


def test_overshoot_of_a_quadratic_lands_on_the_psi_minimiser(orc):
    """A first trial beyond 2 t* (1 - mu) fails sufficient decrease.  psi is a
    quadratic, so the quadratic and the cubic interpolants are exact and the next
    trial is psi's minimiser t* (1 - mu), which passes both conditions."""
    c, t_star = 50.0, 0.04
    phi = quadratic(c, t_star)
    a, sign, evals, loops, reason, calls = orc.mt_search(phi, *phi(0.0), 0.09, 0.1, 0.005)
    assert sign == 1 and evals == 2 and loops == 1 and reason == orc.MT_ACCEPTED
    assert abs(a - t_star * (1 - MU)) <= 1e-12 * t_star
    assert [t for t, _ in calls] == [0.09, a]


def test_planted_dead_loop_fails(orc):
    """The reference's loop start returns the rejected first trial."""
    c, t_star = 50.0, 0.04
    phi = quadratic(c, t_star)
    a, sign, evals, loops, reason, _ = orc.mt_search(phi, *phi(0.0), 0.09, 0.1, 0.005, flags=orc.MT_DEAD_LOOP)
    assert loops == 0 and evals == 1
    assert abs(a - t_star * (1 - MU)) > 1e-3  # the closed form above is missed
    assert not wolfe(phi, a, sign, *phi(0.0))


def test_cubic_overshoot_averages_cubic_and_quadratic_steps(orc):
    """phi = t^3/3 - t from a first trial at 3: psi rises (case 1 of the trial
    selection).  The cubic interpolant of psi is psi itself (minimum sqrt(1 - mu));
    the quadratic through psi(0), psi'(0) and psi(3) is t^2 - (1 - mu) t (minimum
    (1 - mu) / 2).  The cubic step lies farther from 0, so the trial is the mean of
    the two, and it is accepted."""
    a, sign, evals, loops, reason, calls = orc.mt_search(cubic, *cubic(0.0), 3.0, 5.0, 0.001)
    want = 0.5 * ((1 - MU) / 2 + math.sqrt(1 - MU))
    assert reason == orc.MT_ACCEPTED and loops == 1 and evals == 2
    assert abs(a - want) <= 1e-12
    assert wolfe(cubic, a, sign, *cubic(0.0))


def test_cubic_curvature_failure_takes_the_secant_step(orc):
    """From a first trial at 1.6, psi has decreased but phi' = 1.56 > nu: the slopes
    of psi at 0 and 1.6 have opposite signs (case 2).  The trial is whichever of the
    cubic step sqrt(1 - mu) and the secant step (the zero of the line through
    psi'(0) and psi'(1.6), -g_l a_t / (g_t - g_l)) lies farther from 1.6: the secant
    step, which is accepted."""
    a_t = 1.6
    g_l = -(1 - MU)
    g_t = a_t * a_t - 1 + MU
    secant = -g_l * a_t / (g_t - g_l)
    cub = math.sqrt(1 - MU)
    assert abs(cub - a_t) < abs(secant - a_t)
    a, sign, evals, loops, reason, _ = orc.mt_search(cubic, *cubic(0.0), a_t, 5.0, 0.001)
    assert reason == orc.MT_ACCEPTED and loops == 1 and evals == 2
    assert abs(a - secant) <= 1e-12
    assert wolfe(cubic, a, sign, *cubic(0.0))


def test_step_max_clamps_an_acceptable_first_trial(orc):
    """With the minimum beyond step_max the clamped first trial already satisfies
    both conditions: one evaluation, no loop."""
    phi = quadratic(2.0, 0.5)
    a, sign, evals, loops, reason, calls = orc.mt_search(phi, *phi(0.0), 3.0, 0.1, 0.005)
    assert (a, sign, evals, loops, reason) == (0.1, 1, 1, 0, orc.MT_ACCEPTED)
    # step_min clamps from below in the same way
    a, *_ = orc.mt_search(phi, *phi(0.0), 1e-9, 0.1, 0.005)
    assert a == 0.005


def test_ascent_direction_is_reversed(orc):
    """phi'(0) > 0: the search runs along the opposite direction (negative t) and
    returns the same step as for the mirrored function."""
    c, t_star = 50.0, 0.04
    mirrored = quadratic(c, -t_star)
    a, sign, evals, loops, reason, calls = orc.mt_search(mirrored, *mirrored(0.0), 0.09, 0.1, 0.005)
    assert sign == -1 and reason == orc.MT_ACCEPTED and loops == 1
    assert abs(a - t_star * (1 - MU)) <= 1e-12 * t_star
    assert all(t < 0 for t, _ in calls)


def test_zero_slope_takes_no_step(orc):
    phi = quadratic(1.0, 0.0)
    a, sign, evals, loops, reason, calls = orc.mt_search(phi, *phi(0.0), 0.05, 0.1, 0.005)
    assert (a, evals, loops, reason, calls) == (0.0, 0, 0, orc.MT_NO_DESCENT, [])


def test_hessian_only_at_the_accepted_step(orc):
    """With the Hessian kept for the accepted step only, the loop's evaluations ask
    for none and one more evaluation with it follows at the returned step; without
    a loop nothing changes."""
    phi = quadratic(50.0, 0.04)
    a0, *_, calls0 = orc.mt_search(phi, *phi(0.0), 0.09, 0.1, 0.005)
    a1, s1, evals, loops, reason, calls1 = orc.mt_search(phi, *phi(0.0), 0.09, 0.1, 0.005,
                                                        flags=orc.MT_HESSIAN_AT_ACCEPT)
    assert a1 == a0 and evals == 3
    assert calls0 == [(0.09, 1), (a0, 1)]
    assert calls1 == [(0.09, 1), (a0, 0), (a0, 1)]
    clamp = quadratic(2.0, 0.5)
    assert orc.mt_search(clamp, *clamp(0.0), 3.0, 0.1, 0.005, flags=orc.MT_HESSIAN_AT_ACCEPT)[5] == [(0.1, 1)]


# The test functions of More & Thuente (1994), section 5: phi_1 = -a / (a^2 + b),
# phi_2 = (a + b)^5 - 2 (a + b)^4 and phi_3 (a wiggly line).  Their step bounds
# differ from the ones these functions were designed for, so the exits vary; each
# outcome must be one the loop documents.
def mt1(t, b=2.0):
    return -t / (t * t + b), (t * t - b) / (t * t + b) ** 2


def mt2(t, b=0.004):
    u = t + b
    return u ** 5 - 2 * u ** 4, 5 * u ** 4 - 8 * u ** 3


def mt3(t, b=0.01, ell=39):
    if t <= 1 - b:
        p, dp = 1 - t, -1.0
    elif t >= 1 + b:
        p, dp = t - 1, 1.0
    else:
        p, dp = (t - 1) ** 2 / (2 * b) + b / 2, (t - 1) / b
    w = 2 * (1 - b) / (ell * math.pi)
    return p + w * math.sin(ell * math.pi * t / 2), dp + (1 - b) * math.cos(ell * math.pi * t / 2)


@pytest.mark.parametrize("fn", [mt1, mt2, mt3])
@pytest.mark.parametrize("step_init", [1e-3, 1e-1, 1.0, 10.0, 1e3])
def test_more_thuente_functions_end_in_a_documented_exit(orc, fn, step_init):
    phi_0, d_phi_0 = fn(0.0)
    a, sign, evals, loops, reason, calls = orc.mt_search(fn, phi_0, d_phi_0, step_init, 4.0, 1e-3)
    assert 1e-3 <= a <= 4.0 and evals == loops + 1 <= 11
    assert calls[-1][0] == sign * a  # the driver keeps the last evaluation's derivatives
    if reason == orc.MT_ACCEPTED:
        assert wolfe(fn, a, sign, phi_0, d_phi_0)
    elif reason == orc.MT_MAX_ITER:
        assert loops == 10 and not wolfe(fn, a, sign, phi_0, d_phi_0)
    else:
        assert reason == orc.MT_INTERVAL and not wolfe(fn, a, sign, phi_0, d_phi_0)


def test_more_thuente_functions_mostly_accept(orc):
    """Across those starts every search ends with a step that satisfies both
    conditions (the dead loop returns clamped first trials that do not)."""
    live = dead = 0
    for fn in (mt1, mt2, mt3):
        for step_init in (1e-3, 1e-1, 1.0, 10.0, 1e3):
            r = orc.mt_search(fn, *fn(0.0), step_init, 4.0, 1e-3)
            live += r[4] == orc.MT_ACCEPTED
            d = orc.mt_search(fn, *fn(0.0), step_init, 4.0, 1e-3, flags=orc.MT_DEAD_LOOP)
            dead += wolfe(fn, d[0], d[1], *fn(0.0))
    assert live == 15 and dead < live


def test_ndt_hessian_at_accept_keeps_poses(A, orc):
    """At the NDT level, computing the Hessian only at accepted steps (the
    reference's order of passes) changes the pass count by exactly one per Newton
    iteration whose loop ran, and nothing else: the same poses bit for bit."""
    scene = A.scene_cfg(1)
    mp = A.make_map(scene, 60000)
    full = orc.NdtGrid(mp)
    accept = orc.NdtGrid(mp, orc.ndt_params(hessian_at_accept=1))
    scans = [A.make_scan(scene, mp, A.scan_cfg(1500, 1.0, 3.0), 900 + s) for s in range(16)]
    gs = np.stack([s[2] for s in scans])
    r0 = full.align_many([s[0] for s in scans], gs, threads=8)
    r1 = accept.align_many([s[0] for s in scans], gs, threads=8)
    extra = 0
    for x, y in zip(r0, r1):
        assert np.array_equal(x["final_f64"], y["final_f64"])
        assert (x["iterations"], x["converged"]) == (y["iterations"], y["converged"])
        assert 0 <= y["passes"] - x["passes"] <= x["iterations"]
        extra += y["passes"] - x["passes"]
    assert extra > 0  # some loops ran


@pytest.mark.parametrize("hessian_at_accept", [0, 1])
def test_ndt_repeated_trials_keep_poses(A, orc, hessian_at_accept):
    """A line-search trial clamped to the same step as the trial before it (step_min
    or step_max) is at the same pose: taking that pass's sums instead of making the
    pass again gives the same poses, iterations and scores bit for bit, with fewer
    passes; from starts near the optimum most loop trials are such repeats (a Newton
    step shorter than step_min is clamped up to it and overshoots)."""
    scene = A.scene_cfg(1)
    mp = A.make_map(scene, 60000)
    again = orc.NdtGrid(mp, orc.ndt_params(hessian_at_accept=hessian_at_accept, repeat_passes=1))
    reuse = orc.NdtGrid(mp, orc.ndt_params(hessian_at_accept=hessian_at_accept))
    scans = [A.make_scan(scene, mp, A.scan_cfg(1500, 1.0, 3.0), 900 + s) for s in range(12)]
    near = [A.make_scan(scene, mp, A.scan_cfg(1500, 0.01, 0.1), 950 + s) for s in range(12)]
    saved = {}
    for name, group in (("far", scans), ("near", near)):
        gs = np.stack([s[2] for s in group])
        r0 = again.align_many([s[0] for s in group], gs, threads=8)
        r1 = reuse.align_many([s[0] for s in group], gs, threads=8)
        for x, y in zip(r0, r1):
            assert np.array_equal(x["final_f64"], y["final_f64"])
            assert (x["iterations"], x["converged"], x["score"]) == (y["iterations"], y["converged"], y["score"])
            assert y["passes"] <= x["passes"]
        saved[name] = (sum(x["passes"] for x in r0), sum(y["passes"] for y in r1))
    print(saved)
    assert saved["far"][1] < saved["far"][0]
    assert saved["near"][1] < 0.8 * saved["near"][0], saved
